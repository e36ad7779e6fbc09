"""The neighbour-list builds: lanes per atom (1) and one workgroup per cell
(2) over the sorted cell list give rows with the same pairs within the
cutoff in the same order, so forces, energies and trajectories are
bit-identical; the direct build (3: workgroups per cell over the atoms'
cell tags, rows in atom order; a box with fewer than 256 cells deals each
cell's home atoms to up to 8 workgroups) gives the same pairs in another
order, so the same forces to summation order.  Every build meets the FP64 parity bar
against the oracle by itself."""
import numpy as np
import pytest

import lumol_oracle as lo
from lumol_amd import (LennardJones, MolecularDynamics, PairInteraction, Simulation, System, UnitCell,
                       VelocityVerlet)
from lumol_amd.workloads import lj_fluid, nacl, spce

pytestmark = pytest.mark.gpu


def slab():
    # 30 x 30 x 11 A: one cell (nearest image) along z, half cells along x, y
    rng = np.random.default_rng(12)
    n = 700
    pos = rng.uniform(0.0, 1.0, size=(n, 3)) * [30.0, 30.0, 11.0]
    s = System.from_arrays(["Ar"] * n, pos, UnitCell.ortho(30.0, 30.0, 11.0), masses=np.full(n, 39.948))
    s.set_pair_potential(("Ar", "Ar"), PairInteraction(LennardJones(3.4, 1e-4), 5.4))
    return s


def thin_slab():
    # 10.5 x 10.5 x 8 A, rc 3.6: 5 x 5 half cells and one cell along z
    # (nearest image), 25 cells, so the direct build deals each cell's home
    # atoms to 8 workgroups
    rng = np.random.default_rng(14)
    g = np.stack(np.meshgrid(np.arange(6), np.arange(6), np.arange(4), indexing="ij"), -1).reshape(-1, 3)
    pos = (g + 0.5) * [1.75, 1.75, 2.0] + rng.uniform(-0.3, 0.3, size=g.shape)
    n = len(pos)
    s = System.from_arrays(["Ar"] * n, pos, UnitCell.ortho(10.5, 10.5, 8.0), masses=np.full(n, 39.948))
    s.set_pair_potential(("Ar", "Ar"), PairInteraction(LennardJones(1.6, 1e-3), 3.6))
    return s


def clustered():
    # half the atoms in a 9 A corner cube: rows far longer than the mean
    # density predicts, so the first build overflows and the list regrows
    rng = np.random.default_rng(13)
    n = 2000
    pos = np.concatenate([rng.uniform(0.0, 9.0, size=(n // 2, 3)), rng.uniform(0.0, 40.0, size=(n // 2, 3))])
    s = System.from_arrays(["Ar"] * n, pos, UnitCell.cubic(40.0), masses=np.full(n, 39.948))
    s.set_pair_potential(("Ar", "Ar"), PairInteraction(LennardJones(3.4, 1e-4), 8.0))
    return s


CASES = {
    "water-ewald": lambda: spce(nmol=1000, L=31.07, cutoff=7.0, kmax=6),
    "nacl-full-cells": lambda: nacl(ncell=4, cutoff=10.0, kmax=7),
    "lj-fluid": lambda: lj_fluid(4000),
    "slab-nearest-image": slab,
    # 125 half cells: two workgroups per cell in the direct build
    "water-125-cells": lambda: spce(nmol=500, L=24.7, cutoff=8.0, kmax=6),
    "thin-slab-25-cells": thin_slab,
    "clustered-overflow": clustered,
}


def evaluate(make, build):
    s = make()
    s.set_tuning(nlist_build=build)
    f = s.forces()
    e = s.potential_energy()
    v = s.virial()
    assert s.tuning()["nlist_build"] == build
    return s, f, e, v


@pytest.mark.parametrize("case", list(CASES))
def test_builds_give_identical_forces(case):
    _, f1, e1, v1 = evaluate(CASES[case], 1)
    s2, f2, e2, v2 = evaluate(CASES[case], 2)
    assert np.array_equal(f1, f2)
    assert e1 == e2
    assert np.array_equal(v1, v2)
    fo = lo.OracleSystem(s2, nthreads=8).forces()
    assert np.abs(f2 - fo).max() <= 1e-10 * np.abs(fo).max()


@pytest.mark.parametrize("case", list(CASES))
def test_direct_build_matches(case):
    s2, f2, e2, v2 = evaluate(CASES[case], 2)
    s3, f3, e3, v3 = evaluate(CASES[case], 3)
    assert np.abs(f3 - f2).max() <= 1e-12 * np.abs(f2).max()
    assert abs(e3 - e2) <= 1e-12 * abs(e2)
    assert np.abs(v3 - v2).max() <= 1e-12 * np.abs(v2).max()
    fo = lo.OracleSystem(s3, nthreads=8).forces()
    assert np.abs(f3 - fo).max() <= 1e-10 * np.abs(fo).max()


@pytest.mark.parametrize("case", ["water-ewald", "water-125-cells", "lj-fluid"])
def test_direct_build_trajectory(case):
    # 60 steps: the direct build's list is rebuilt from the tags the first
    # half-steps write, and stays within summation-order noise of build 2
    out = []
    for build in (2, 3):
        s = CASES[case]()
        s.set_tuning(nlist_build=build)
        Simulation(MolecularDynamics(VelocityVerlet(0.5 if case.startswith("water") else 2.0))).run(s, 60)
        out.append(s.positions.copy())
    assert np.abs(out[0] - out[1]).max() < 1e-8


def test_direct_build_mixed_and_positions_reset():
    # mixed form, then new positions from the host: the next evaluation's
    # displacement test writes the tags the rebuild reads
    from lumol_amd._native import LG_PRECISION_MIXED
    s = spce(nmol=1000, L=31.07, cutoff=7.0, kmax=6)
    s.set_precision(LG_PRECISION_MIXED)
    s.set_tuning(nlist_build=3)
    s.forces()
    rng = np.random.default_rng(4)
    pos = s.positions + rng.uniform(-0.5, 0.5, size=s.positions.shape)
    s.positions = pos
    f = s.forces()
    assert s.tuning()["nlist_build"] == 3
    fo = lo.OracleSystem(s, nthreads=8).forces()
    rms = np.sqrt(((f - fo) ** 2).sum() / (fo ** 2).sum())
    assert rms < 1e-5


@pytest.mark.parametrize("case", ["water-ewald", "lj-fluid"])
def test_builds_give_identical_trajectories(case):
    out = []
    for build in (1, 2):
        s = CASES[case]()
        s.set_tuning(nlist_build=build)
        Simulation(MolecularDynamics(VelocityVerlet(0.5 if case == "water-ewald" else 2.0))).run(s, 80)
        out.append((s.positions.copy(), s.velocities.copy()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_default_choice():
    big = spce(nmol=3000, L=44.8, cutoff=10.0, kmax=4)  # 512 half cells, 17 atoms each
    assert big.tuning()["nlist_build"] == 3
    small = nacl(ncell=3, cutoff=8.0, kmax=5)
    assert small.tuning()["nlist_build"] == 1
    # past 32768 atoms the per-cell build keeps the cell list
    huge = lj_fluid(40000)
    huge.set_tuning(nlist_build=3)
    assert huge.tuning()["nlist_build"] == 2
