"""GPU parity: the FP64 HIP path against the oracle on identical inputs.

Bar (north star): forces, energies and virial to 1e-10 relative; MD
positions to 1e-8 A after 100 steps.  Relative force error is measured
per atom against the largest force component of the system (`frel`), and
energies/virials against their own magnitude.  The pair set of the device
neighbour search must equal the brute-force cutoff loop's exactly.
"""
import numpy as np
import pytest

import lumol_oracle as lo
import reference_values as R
from helpers import (nist_lj_system, nist_spce_system, random_molecular_system, read_lammps_forces,
                     water_molecule)
from lumol_amd import (BerendsenThermostat, Ewald, LeapFrog, MolecularDynamics, PairRestriction, Simulation,
                       System, UnitCell, VelocityVerlet, Wolf)
from lumol_amd.workloads import argon, nacl, spce

pytestmark = pytest.mark.gpu

TOL = 1e-10


def frel(f, g):
    scale = np.abs(g).max()
    return np.abs(f - g).max() / scale


def erel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def wrel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def check_system(s, nthreads=8, tol=TOL, force_floor=0.0):
    """Forces, energy terms and virial against the oracle at `tol`.  Forces
    are measured against the largest force component, or `force_floor` when
    that is larger (a perfect crystal or an ideal gas has forces that are
    rounding noise on both sides)."""
    o = lo.OracleSystem(s, nthreads=nthreads)
    fg, fo = s.forces(), o.forces()
    fscale = max(np.abs(fo).max(), force_floor)
    assert np.abs(fg - fo).max() <= tol * fscale, f"forces rel {np.abs(fg - fo).max() / fscale:.3e}"
    tg, to = s.energy_terms(), o.energy_terms()
    # each term to `tol` relative; a term that is rounding noise next to the
    # total (e.g. angles at their equilibrium) is held to tol * |total|
    scale = max(abs(to[k]) for k in ("pairs", "pairs_tail", "bonds", "angles", "coulomb"))
    for key in ("pairs", "pairs_tail", "bonds", "angles", "coulomb"):
        a, b = getattr(tg, key), to[key]
        assert abs(a - b) <= tol * max(abs(b), 1e-6 * scale), f"{key} {a} {b}"
    assert erel(tg.total, to["total"]) < tol
    if not s.cell.is_infinite():
        wg, wo = s.virial(), o.virial()
        assert wrel(wg, wo) < tol, f"virial rel {wrel(wg, wo):.3e}"
    return o


# --- unit systems -----------------------------------------------------------------------
def test_wolf_two_ions():
    d = R.WOLF_DOC
    s = System.from_arrays(["Na", "Cl"], [(0, 0, 0), (2, 0, 0)], UnitCell.cubic(d["cell"]), charges=[1.0, -1.0])
    s.set_coulomb_potential(Wolf(d["cutoff"]))
    assert erel(s.potential_energy(), d["energy"]) < 1e-14


def test_ewald_water_molecule():
    s = water_molecule()
    ew = Ewald(8.0, 10)
    ew.set_restriction(PairRestriction.INTER_MOLECULAR)
    s.set_coulomb_potential(ew)
    # the total is a 1e-5 remainder of 0.1-sized real, self and k-space terms
    assert erel(s.potential_energy(), R.EWALD_WATER["energy"]) < TOL


def test_ewald_nacl_pair():
    d = R.EWALD_NACL_PAIR
    s = System.from_arrays(["Cl", "Na"], [(0, 0, 0), (1.5, 0, 0)], UnitCell.cubic(d["cell"]), charges=[-1.0, 1.0])
    s.set_coulomb_potential(Ewald(d["cutoff"], d["kmax"]))
    check_system(s)


@pytest.mark.parametrize("cutoff", [3.0, 4.0])
@pytest.mark.parametrize("name", ["lj-1.xyz", "lj-2.xyz", "lj-3.xyz", "lj-4.xyz"])
def test_nist_lj(name, cutoff):
    (e_ref, e_dec), (w_ref, w_dec), (t_ref, t_dec) = R.NIST_LJ[cutoff][name]
    s = nist_lj_system(name, cutoff, tail=False)
    energy = s.potential_energy()
    assert round(energy, e_dec) == e_ref
    assert round(np.trace(s.virial()), w_dec) == w_ref
    st = nist_lj_system(name, cutoff, tail=True)
    assert round(st.potential_energy() - energy, t_dec) == t_ref
    check_system(st)


@pytest.mark.parametrize("cutoff", [9.0, 10.0])
@pytest.mark.parametrize("name", ["spce-1.xyz", "spce-2.xyz", "spce-3.xyz", "spce-4.xyz"])
def test_nist_spce(name, cutoff):
    total, pairs, tail, coul = R.NIST_SPCE_ENERGIES[cutoff][name]
    s = nist_spce_system(name, cutoff)
    t = s.energy_terms()
    kb = R.K_BOLTZMANN
    assert erel(t.total / kb, total) < 1e-3
    assert erel(t.coulomb / kb, coul) < 1e-3
    check_system(s)


@pytest.mark.parametrize("cutoff", [9.0, 10.0])
@pytest.mark.parametrize("name", ["spce-1.xyz", "spce-2.xyz"])
def test_nist_spce_lammps_forces(name, cutoff):
    kmax, alpha, ffile = R.NIST_SPCE_LAMMPS[cutoff][name]
    s = nist_spce_system(name, cutoff, kmax=kmax, alpha=alpha, tail=False)
    f = s.forces() / R.KCAL_PER_MOL
    expected = read_lammps_forces(ffile)
    assert (np.abs((f - expected) / expected) < 5e-3).all()


@pytest.mark.parametrize("coulomb", ["ewald", "wolf", None])
@pytest.mark.parametrize("restriction", [PairRestriction.INTER_MOLECULAR, PairRestriction.EXCLUDE13,
                                         PairRestriction.NONE])
def test_molecular_restrictions(coulomb, restriction):
    s = random_molecular_system(n_mol=30, L=15.0, coulomb=coulomb, restriction=restriction)
    check_system(s)


# --- benchmark configurations at oracle-friendly sizes -----------------------------------------
def test_argon_c1():
    check_system(argon())


def test_nacl_small():
    check_system(nacl(ncell=4, cutoff=10.0, kmax=7))


def test_spce_small():
    check_system(spce(nmol=300, L=20.8, cutoff=8.0, kmax=7))


def test_nacl_c2_full_size_forces():
    s = nacl()
    o = lo.OracleSystem(s, nthreads=16)
    assert frel(s.forces(), o.forces()) < TOL


# --- pair set ---------------------------------------------------------------------------------
@pytest.mark.parametrize("builder", ["argon", "nacl", "random"])
def test_pair_set_bit_exact(builder):
    if builder == "argon":
        s, rc = argon(), 8.5
    elif builder == "nacl":
        s, rc = nacl(ncell=4, cutoff=10.0, kmax=7), 10.0
    else:
        rng = np.random.default_rng(5)
        n, L = 2000, 30.0
        s = System.from_arrays(["Ar"] * n, rng.uniform(-2 * L, 2 * L, size=(n, 3)), UnitCell.cubic(L))
        from lumol_amd import LennardJones, PairInteraction
        s.set_pair_potential(("Ar", "Ar"), PairInteraction(LennardJones(3.4, 1.0), 9.0))
        rc = 9.0
    got = s.pairs_within(rc)
    want = lo.OracleSystem(s).pairs_within(rc)
    assert got.shape == want.shape and (got == want).all()


# --- molecular dynamics -------------------------------------------------------------------------
def md_compare(s, sim, oracle_integrator, dt, nsteps, bT=-1.0, tau=1.0, tol=1e-8):
    o = lo.OracleSystem(s, nthreads=16)
    o.md_reset()
    sim.run(s, nsteps)
    o.md_run(oracle_integrator, dt, nsteps, bT, tau)
    dp = np.abs(s.positions - o.positions()).max()
    assert dp < tol, f"max position deviation {dp:.3e} A"
    dv = np.abs(s.velocities - o.velocities()).max() / np.abs(o.velocities()).max()
    assert dv < 1e-6


def test_md_argon_vv_100_steps():
    s = argon()
    md_compare(s, Simulation(MolecularDynamics(VelocityVerlet(1.0))), 0, 1.0, 100)


def test_md_argon_leapfrog_100_steps():
    s = argon()
    md_compare(s, Simulation(MolecularDynamics(LeapFrog(1.0))), 1, 1.0, 100)


def test_md_nacl_berendsen_100_steps():
    s = nacl(ncell=4, cutoff=10.0, kmax=7)
    md = MolecularDynamics(VelocityVerlet(1.0))
    md.set_thermostat(BerendsenThermostat(1200.0, 100.0))
    md_compare(s, Simulation(md), 0, 1.0, 100, 1200.0, 100.0)


def test_md_water_100_steps():
    s = spce(nmol=216, L=18.64, cutoff=8.0, kmax=6)
    md_compare(s, Simulation(MolecularDynamics(VelocityVerlet(0.5))), 0, 0.5, 100)


def test_md_deterministic():
    a, b = nacl(ncell=4, cutoff=10.0, kmax=7), nacl(ncell=4, cutoff=10.0, kmax=7)
    for s in (a, b):
        Simulation(MolecularDynamics(VelocityVerlet(1.0))).run(s, 50)
    assert (a.positions == b.positions).all() and (a.velocities == b.velocities).all()


def test_nve_energy_conservation_nacl_c2():
    # size-independent property at the full C2 size: NVE drift bound
    s = nacl()
    e0 = s.total_energy()
    Simulation(MolecularDynamics(VelocityVerlet(1.0))).run(s, 200)
    e1 = s.total_energy()
    assert abs((e1 - e0) / e0) < 5e-4
    f = s.forces()
    assert np.abs(f.sum(0)).max() < 1e-9 * np.abs(f).max() * np.sqrt(len(f))
