"""GPU parity of the potential kinds beyond LJ/Harmonic/Null and of
TableComputation (functions.rs:162-592, computations.rs:70-175): the
general pair-kernel instance, the intra-molecular kernel with Morse,
CosineHarmonic and Torsion bonded terms, the Monte Carlo kernels and MD, all
against the oracle.

Bars as in test_gpu_parity: forces, energies and virial to 1e-10 relative
(FP64 form); 1e-8 A after 100 MD steps; Monte Carlo costs as test_gpu_mc.
In the mixed form these systems run the FP64 pair kernel (only the k-space
is FP32), so the mixed form is held to the mixed-form bar of the k-space.
"""
import numpy as np
import pytest

import lumol_oracle as lo
from helpers import other_potentials_system
from lumol_amd import MolecularDynamics, PairRestriction, Simulation, VelocityVerlet
from lumol_amd import _native as nat
from lumol_amd._native import LG_PRECISION_MIXED
from test_gpu_abi_errors import new_system, status_and_message
from test_gpu_mc import check_moves
from test_gpu_parity import check_system, md_compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("coulomb", ["wolf", "ewald", None])
@pytest.mark.parametrize("variant", [1, 2])
def test_other_potentials_fp64(variant, coulomb):
    s = other_potentials_system(variant=variant, coulomb=coulomb)
    o = check_system(s)
    t = o.energy_terms()
    assert t["pairs"] != 0.0 and t["pairs_tail"] != 0.0 and t["angles"] != 0.0 and t["dihedrals"] != 0.0


@pytest.mark.parametrize("restriction", [PairRestriction.NONE, PairRestriction.scale14(0.5),
                                         PairRestriction.INTER_MOLECULAR], ids=["none", "scale14", "inter"])
def test_other_potentials_restrictions(restriction):
    # 1-2 and 1-3 pairs (restriction none) meet the potentials at short range
    s = other_potentials_system(variant=2, coulomb="wolf", restriction=restriction)
    check_system(s)


def test_other_potentials_mixed_form():
    s = other_potentials_system(variant=1, coulomb="ewald", precision=LG_PRECISION_MIXED)
    o = lo.OracleSystem(s, nthreads=8)
    fg, fo = s.forces(), o.forces()
    err = np.sqrt(((fg - fo) ** 2).sum() / (fo ** 2).sum())
    assert err < 1e-5, err
    assert abs(s.potential_energy() - o.potential_energy()) <= 1e-5 * abs(o.potential_energy())


@pytest.mark.parametrize("variant", [1, 2])
def test_other_potentials_md_100_steps(variant):
    s = other_potentials_system(variant=variant, coulomb="wolf", morse=False)
    rng = np.random.default_rng(3)
    s.velocities = rng.normal(scale=2e-3, size=(s.size(), 3))
    md_compare(s, Simulation(MolecularDynamics(VelocityVerlet(1.0))), 0, 1.0, 100)


@pytest.mark.parametrize("variant", [1, 2])
def test_other_potentials_mc_moves(variant):
    s = other_potentials_system(variant=variant, coulomb="wolf")
    check_moves(s, nmoves=24, amplitude=0.5)


def test_table_matches_oracle_at_the_grid_end():
    # Two atoms at distances around the table's last bins, whose cutoff
    # reaches past the table: from r >= max - delta the table gives 0
    # (computations.rs:127-133)
    from lumol_amd import LennardJones, PairInteraction, System, TableComputation, UnitCell
    size, mx = 100, 5.0
    delta = mx / size
    for x in (4.0, mx - 2 * delta + 1e-3, mx - delta - 1e-9, mx - delta, mx - 0.5 * delta, 5.5):
        s = System.from_arrays(["Ar", "Ar"], [(0.0, 0.0, 0.0), (x, 0.0, 0.0)], UnitCell.cubic(20.0))
        s.set_pair_potential(("Ar", "Ar"),
                             PairInteraction(TableComputation(LennardJones(2.0, 0.01), size, mx), 6.0))
        o = lo.OracleSystem(s)
        assert np.array_equal(s.forces() == 0.0, o.forces() == 0.0), x
        if x < mx - delta:
            check_system(s)
        else:
            assert s.potential_energy() == 0.0 and o.potential_energy() == 0.0


def test_potential_kind_errors():
    L = nat.lib()
    h = new_system()
    g = np.array([8.0, -2.0])
    rc, msg = status_and_message(L.lg_pair_set(h, 0, 0, nat.LG_POT_GAUSSIAN, nat.dptr(g), 5.0, 0, 0, 0, 1.0))
    assert rc == nat.LG_E_INVALID and "positive" in msg
    mie = np.array([2.0, 1.0, 6.0, 12.0])
    rc, msg = status_and_message(L.lg_pair_set(h, 0, 0, nat.LG_POT_MIE, nat.dptr(mie), 5.0, 0, 0, 0, 1.0))
    assert rc == nat.LG_E_INVALID and "exponent" in msg
    ch = np.array([1.0, 2.0])
    rc, msg = status_and_message(L.lg_pair_set(h, 0, 0, nat.LG_POT_COSINE_HARMONIC, nat.dptr(ch), 5.0, 0, 0, 0,
                                               1.0))
    assert rc == nat.LG_E_UNSUPPORTED and "not a pair potential" in msg
    tor = np.array([1.0, 0.0, 3.0])
    assert L.lg_angle_set(h, 0, 0, 0, nat.LG_POT_TORSION, nat.dptr(tor)) == nat.LG_E_UNSUPPORTED
    assert L.lg_bond_set(h, 0, 0, nat.LG_POT_LJ, nat.dptr(ch)) == nat.LG_E_UNSUPPORTED
    assert L.lg_dihedral_set(h, 0, 0, 0, 0, nat.LG_POT_TORSION, nat.dptr(np.array([1.0, 0.0, 2.5]))) == \
        nat.LG_E_INVALID
    assert L.lg_dihedral_set(h, 0, 0, 0, 0, nat.LG_POT_TORSION, nat.dptr(tor)) == nat.LG_OK
    lj = np.array([3.4, 1e-4])
    assert L.lg_pair_set_table(h, 0, 0, nat.LG_POT_LJ, nat.dptr(lj), 0, 5.0, 5.0, 0, 0, 0, 1.0) == nat.LG_E_INVALID
    assert L.lg_pair_set_table(h, 0, 0, nat.LG_POT_LJ, nat.dptr(lj), 100, -1.0, 5.0, 0, 0, 0, 1.0) == \
        nat.LG_E_INVALID
    assert L.lg_pair_set_table(h, 0, 0, nat.LG_POT_LJ, nat.dptr(lj), 100, 5.0, 5.0, 1, 1, 0, 1.0) == nat.LG_OK
    f = np.zeros(6)
    assert L.lg_forces(h, nat.dptr(f), 0) == nat.LG_OK
    assert np.isfinite(f).all()
    L.lg_system_destroy(h)
