"""GPU parity of the Monte Carlo path (EnergyCache + Translate + Metropolis)
against the oracle's restatement of lumol-core/src/sys/cache.rs and
lumol-sim/src/mc.

* Move costs: the reference's cache tests compare cached deltas at rel 1e-9
  (cache.rs:397-484); the device recomputes the old pair energies instead of
  caching them, so only the summation order differs: rel 1e-9, with an
  absolute floor of 1e-12 |E| for costs that cancel to ~0.
* Chains: both sides draw the same Philox4x32-10 numbers per attempt, so
  accept/reject decisions, accepted counts and final positions must be
  identical; the running cache energy to 1e-10 relative.
"""
import numpy as np
import pytest

import lumol_oracle as lo
from lumol_amd import PairRestriction, System, UnitCell, Wolf, mc_run_batch
from lumol_amd.workloads import lj_mc_chain, mc_translate, nacl, spce

pytestmark = pytest.mark.gpu


def cost_close(a, b, energy):
    return abs(a - b) <= max(1e-9 * abs(b), 1e-12 * abs(energy))


def random_moves(s, rng, nmoves, amplitude):
    mols = s.molecules()
    for _ in range(nmoves):
        m = int(rng.integers(len(mols)))
        a, b = mols[m]
        disp = rng.uniform(-amplitude, amplitude, size=3)
        yield m, s.positions[a:b] + disp


def check_moves(s, nmoves=24, amplitude=0.6, seed=11):
    o = lo.OracleSystem(s, nthreads=8)
    e_gpu, e_ref = s.cache_init(), None
    o.cache_init()
    e_ref = o.cache_energy()
    assert abs(e_gpu - e_ref) <= 1e-10 * abs(e_ref)
    rng = np.random.default_rng(seed)
    for k, (m, newpos) in enumerate(random_moves(s, rng, nmoves, amplitude)):
        cg, co = s.move_cost(m, newpos), o.move_cost(m, newpos)
        assert cost_close(cg, co, e_ref), f"move {k}: {cg!r} vs {co!r}"
        if k % 2 == 0:
            s.move_commit(True)
            o.move_commit(True)
        else:
            s.move_discard()
            o.move_discard()
    assert np.array_equal(s.positions, o.positions())
    eg, eo = s.cache_energy(), o.cache_energy()
    assert abs(eg - eo) <= 1e-10 * abs(eo)
    # the running cache agrees with a fresh evaluation of the new configuration
    assert abs(eg - s.potential_energy()) <= 1e-9 * abs(eg)


def test_move_cost_lj():
    s, _, _ = lj_mc_chain(500, seed=7)
    check_moves(s)


def test_move_cost_water_ewald():
    s = spce(nmol=100, L=14.4, cutoff=7.0, kmax=5)
    check_moves(s, amplitude=0.4)


def test_move_cost_nacl_wolf():
    s = nacl(ncell=3, cutoff=8.0, kmax=5)
    s.set_coulomb_potential(Wolf(8.0))
    check_moves(s, amplitude=0.3)


def test_commit_without_cost_is_an_error():
    s, _, _ = lj_mc_chain(108)
    s.cache_init()
    with pytest.raises(Exception):
        s.move_commit(True)


def run_both(s, seed, attempts, beta, delta, attempt0=0):
    o = lo.OracleSystem(s, nthreads=8)
    o.cache_init()
    s.cache_init()
    acc_g = s.mc_run(seed, attempt0, attempts, beta, delta)
    acc_o = o.mc_run(seed, attempt0, attempts, beta, delta)
    return o, acc_g, acc_o


@pytest.mark.parametrize("natoms", [108, 500, 1000])
def test_mc_chain_matches_oracle(natoms):
    # 108 and 500 atoms: fewer than 4 cells per axis (all-atoms loop); 1000:
    # 4 cells per axis (cell list)
    s, beta, delta = lj_mc_chain(natoms, seed=natoms)
    o, acc_g, acc_o = run_both(s, seed=1234 + natoms, attempts=3000, beta=beta, delta=delta)
    assert acc_g == acc_o and 0 < acc_g < 3000
    assert np.array_equal(s.positions, o.positions())
    eg, eo = s.cache_energy(), o.cache_energy()
    assert abs(eg - eo) <= 1e-10 * abs(eo)
    assert abs(eg - s.potential_energy()) <= 1e-9 * abs(eg)


def test_mc_chain_water_ewald():
    # 4 cells per axis: molecules and Ewald through the cell-list path
    s = spce(nmol=512, L=24.86, cutoff=6.0, kmax=5)
    beta = 1.0 / (lo_kb() * 300.0)
    o, acc_g, acc_o = run_both(s, seed=77, attempts=400, beta=beta, delta=0.5)
    assert acc_g == acc_o and acc_g > 0
    assert np.array_equal(s.positions, o.positions())
    eg, eo = s.cache_energy(), o.cache_energy()
    assert abs(eg - eo) <= 1e-10 * abs(eo)


def lo_kb():
    from lumol_amd import K_BOLTZMANN
    return K_BOLTZMANN


def test_mc_resume_equals_single_run():
    a, beta, delta = lj_mc_chain(1000, seed=3)
    b, _, _ = lj_mc_chain(1000, seed=3)
    a.mc_run(9, 0, 2000, beta, delta)
    b.mc_run(9, 0, 1000, beta, delta)
    b.mc_run(9, 1000, 1000, beta, delta)
    assert np.array_equal(a.positions, b.positions)
    assert a.cache_energy() == b.cache_energy() or abs(a.cache_energy() - b.cache_energy()) <= 1e-12 * abs(
        a.cache_energy())


def test_mc_batch_matches_single_chains():
    sizes = [108, 1000, 500, 1200]
    chains = [lj_mc_chain(n, seed=20 + n) for n in sizes]
    systems = [c[0] for c in chains]
    oracles = []
    for (s, beta, delta), n in zip(chains, sizes):
        o = lo.OracleSystem(s, nthreads=8)
        o.cache_init()
        oracles.append((o, o.mc_run(100 + n, 0, 1500, beta, delta)))
    acc = mc_run_batch(systems, [100 + n for n in sizes], 0, 1500, [c[1] for c in chains], [c[2] for c in chains])
    for s, (o, acc_o), a in zip(systems, oracles, acc):
        assert int(a) == acc_o
        assert np.array_equal(s.positions, o.positions())


def test_mc_cell_overflow_regrows():
    # 120 atoms packed in one corner of a 41 A box with 10 A cutoff (4 cells
    # per axis): the first binning exceeds the default cell capacity and is
    # redone with room for every atom; the chain must still follow the
    # oracle exactly
    rng = np.random.default_rng(5)
    n = 120
    pos = rng.uniform(0.0, 6.0, size=(n, 3))
    s = System.from_arrays(["Ar"] * n, pos, UnitCell.cubic(41.0), masses=np.full(n, 39.948))
    from lumol_amd import LennardJones, PairInteraction
    s.set_pair_potential(("Ar", "Ar"), PairInteraction(LennardJones(0.4, 1e-4), 10.0))
    o, acc_g, acc_o = run_both(s, seed=5, attempts=800, beta=1.0 / (lo_kb() * 300.0), delta=1.0)
    assert acc_g == acc_o
    assert np.array_equal(s.positions, o.positions())


def test_mc_simulation_api():
    s, beta, delta = lj_mc_chain(108, seed=8)
    temperature = 1.0 / (lo_kb() * beta)
    sim = mc_translate(temperature, delta, seed=42)
    sim.run(s, 500)
    sim.run(s, 500)
    t, _, _ = lj_mc_chain(108, seed=8)
    t.cache_init()
    t.mc_run(42, 0, 1000, beta, delta)
    assert np.array_equal(s.positions, t.positions)
    assert sim.propagator.attempts == 1000
