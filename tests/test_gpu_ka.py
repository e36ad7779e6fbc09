"""Kob-Andersen binary LJ mixture (the 1,000,000-atom configuration's
model) at oracle-friendly sizes, and size-independent properties at a size
that uses the same launch shapes as 1M atoms."""
import numpy as np
import pytest

import lumol_oracle as lo
from lumol_amd import MolecularDynamics, Simulation, VelocityVerlet
from lumol_amd._native import LG_PRECISION_MIXED
from lumol_amd.workloads import kob_andersen

pytestmark = pytest.mark.gpu


def test_ka_fp64_and_mixed_parity():
    s = kob_andersen(2000)
    o = lo.OracleSystem(s, nthreads=8)
    fo, eo = o.forces(), o.potential_energy()
    f = s.forces()
    assert np.abs(f - fo).max() / np.abs(fo).max() < 1e-10
    assert abs(s.potential_energy() - eo) < 1e-10 * abs(eo)
    s.set_precision(LG_PRECISION_MIXED)
    d = s.forces() - fo
    assert np.sqrt((d * d).sum() / (fo * fo).sum()) < 1e-5
    assert abs(s.potential_energy() - eo) < 1e-4 * abs(eo)


def test_ka_large_momentum_and_drift():
    # 131072 atoms: 4 lanes per atom as at 1M atoms.  The KA potential is
    # plain LJ cut at 2.5 sigma (unshifted), so pairs crossing the cutoff
    # change the total energy by V(rc) each: energy is not a constant of the
    # model here, and the bound is relative to the FP64 form (mixed drift
    # within 2x the FP64 drift, same final energy to 1e-4).
    drift, final = {}, {}
    for form in (0, LG_PRECISION_MIXED):
        s = kob_andersen(131072, seed=8)
        s.set_precision(form)
        sim = Simulation(MolecularDynamics(VelocityVerlet(1.0)))
        sim.run(s, 50)  # leave the lattice start
        f = s.forces()
        assert np.abs(f.sum(0)).max() < 1e-6 * np.abs(f).max() * np.sqrt(len(f))
        e0 = s.total_energy()
        sim.run(s, 200)
        final[form] = s.total_energy()
        drift[form] = abs(final[form] - e0) / abs(e0)
        s.close()
    assert drift[LG_PRECISION_MIXED] <= 2.0 * drift[0] + 1e-6
    assert abs(final[LG_PRECISION_MIXED] - final[0]) <= 1e-4 * abs(final[0])


@pytest.mark.parametrize("build", [1, 2])
def test_ka_cell_list_path_parity(build):
    # 40000 atoms: past the single-workgroup cell list and the direct build,
    # so the four-kernel cell list (count, scan, scatter, rank) feeds the
    # Verlet build; both builds give the same rows and meet the oracle
    s = kob_andersen(40000, seed=5)
    s.set_tuning(nlist_build=build)
    assert s.tuning()["nlist_build"] == build
    f = s.forces()
    fo = lo.OracleSystem(s, nthreads=8).forces()
    assert np.abs(f - fo).max() / np.abs(fo).max() < 1e-10
    t = kob_andersen(40000, seed=5)
    t.set_tuning(nlist_build=3 - build)
    assert np.array_equal(t.forces(), f)


def test_ka_one_million_atoms_properties():
    # the configuration's own size: 1,000,000 atoms.  Size-independent
    # properties: total force ~ 0 (Newton's third law over the full list),
    # the mixed form's forces against the FP64 form's (RMS < 1e-5), and the
    # energy change over 100 steps of the mixed form within twice the FP64
    # form's (plus 1e-6), both starting from the same equilibrated state.
    s = kob_andersen(1_000_000)
    sim = Simulation(MolecularDynamics(VelocityVerlet(1.0)))
    sim.run(s, 20)  # leave the lattice start
    pos, vel = s.positions.copy(), s.velocities.copy()
    f64 = s.forces()
    assert np.abs(f64.sum(0)).max() < 1e-6 * np.abs(f64).max() * np.sqrt(len(f64))
    change = {}
    for form in (0, LG_PRECISION_MIXED):
        s.set_precision(form)
        s.positions, s.velocities = pos, vel
        if form:
            d = s.forces() - f64
            assert np.sqrt((d * d).sum() / (f64 * f64).sum()) < 1e-5
        e0 = s.total_energy()
        run = Simulation(MolecularDynamics(VelocityVerlet(1.0)))
        run.run(s, 100)
        e1 = s.total_energy()
        assert np.isfinite(e1)
        change[form] = abs(e1 - e0) / abs(e0)
    s.close()
    print(f"KA 1M: relative energy change over 100 steps fp64 {change[0]:.2e}, "
          f"mixed {change[LG_PRECISION_MIXED]:.2e}")
    assert change[LG_PRECISION_MIXED] <= 2.0 * change[0] + 1e-6
