"""The two neighbour-list builds (lanes per atom; one workgroup per cell with
candidates staged in LDS) give rows with the same pairs within the cutoff in
the same order, so forces, energies and trajectories are bit-identical; the
cell build also meets the FP64 parity bar against the oracle by itself."""
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
    assert big.tuning()["nlist_build"] == 2
    small = nacl(ncell=3, cutoff=8.0, kmax=5)
    assert small.tuning()["nlist_build"] == 1
