"""Launch knobs and the autotuner: every setting gives the FP64 parity bar
against the oracle, and the autotuner's choice changes forces only by
summation order."""
import numpy as np
import pytest

import lumol_oracle as lo
from lumol_amd import tuning
from lumol_amd.workloads import nacl, spce

pytestmark = pytest.mark.gpu


def frel(f, g):
    return np.abs(f - g).max() / np.abs(g).max()


@pytest.mark.parametrize("lanes", [4, 8, 16, 32, 64])
@pytest.mark.parametrize("split", [1, 3, 32])
def test_every_knob_setting_meets_parity(lanes, split):
    s = nacl(ncell=4, cutoff=10.0, kmax=7)
    fo = lo.OracleSystem(s, nthreads=8).forces()
    kind = 1 + (lanes + split) % 3
    s.set_tuning(pair_lanes=lanes, kspace_split=split, rho_width=(64, 128, 256)[lanes % 3], rho_kind=kind)
    assert frel(s.forces(), fo) < 1e-10
    t = s.tuning()
    assert t["pair_lanes"] == lanes and t["kspace_split"] == split and t["rho_width"] == (64, 128, 256)[lanes % 3]
    assert t["rho_kind"] == kind


def test_autotune_choice_and_cache():
    s = spce(nmol=300, L=20.8, cutoff=8.0, kmax=7)
    f0 = s.forces()
    best = tuning.autotune(s, reps=3)
    assert best["pair_lanes"] in tuning.PAIR_LANES
    assert best["kspace_split"] in tuning.KSPACE_SPLITS and best["rho_kind"] in (1, 2, 3)
    assert best["rho_width"] in tuning.RHO_WIDTHS if best["rho_kind"] != 2 else best["rho_width"] == 0
    assert frel(s.forces(), f0) < 1e-12
    # a second system of the same bucket reuses the choice
    t = spce(nmol=300, L=20.8, cutoff=8.0, kmax=7, seed=9)
    assert tuning.autotune(t) == best
    assert tuning.bucket(t) in tuning.cached()


def test_autotune_on_md_steps():
    # candidates timed on MD steps (the launch mix of lg_md_run); the choice
    # changes forces only by summation order
    from lumol_amd import MolecularDynamics, Simulation, VelocityVerlet
    from lumol_amd._native import LG_PRECISION_MIXED
    s = spce(nmol=300, L=20.8, cutoff=8.0, kmax=7, seed=12)
    s.set_precision(LG_PRECISION_MIXED)
    sim = Simulation(MolecularDynamics(VelocityVerlet(0.5)))
    best = tuning.autotune(s, use_cache=False, rounds=1, step=lambda: sim.run(s, 5))
    assert best["pair_lanes"] in tuning.PAIR_LANES and best["kspace_split"] in tuning.KSPACE_SPLITS
    assert best["rho_kind"] in (1, 2, 3)
    t = spce(nmol=300, L=20.8, cutoff=8.0, kmax=7, seed=12)
    t.set_precision(LG_PRECISION_MIXED)
    f0 = t.forces()
    t.set_tuning(best["pair_lanes"], best["kspace_split"], best["rho_width"], rho_kind=best["rho_kind"])
    # the mixed form's sums differ by FP32 rounding between launch shapes
    assert frel(t.forces(), f0) < 1e-5
    assert np.isfinite(s.positions).all()


def test_autotune_concurrent_applies_one_choice():
    # two water systems on their own streams: one set of knobs for both,
    # chosen by timing them together; forces change only by summation order
    import threading
    import time

    from lumol_amd import MolecularDynamics, Simulation, VelocityVerlet
    systems = [spce(nmol=300, L=20.8, cutoff=8.0, kmax=7, seed=s) for s in (21, 22)]
    for k, s in enumerate(systems):
        s.stream_slot = 1 + k
    f0 = [s.forces() for s in systems]
    calls = []

    def run():
        def work(s):
            Simulation(MolecularDynamics(VelocityVerlet(0.5))).run(s, 5)
        ts = [threading.Thread(target=work, args=(s,)) for s in systems]
        t0 = time.perf_counter()
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        calls.append(1)
        return time.perf_counter() - t0

    best = tuning.autotune_concurrent(systems, run, rounds=1, passes=1)
    assert len(calls) == (len(tuning.PAIR_LANES) + len(tuning.KSPACE_SPLITS) + len(tuning.RHO_WIDTHS)
                          + len(tuning.RHO_KINDS) + len(tuning.NLIST_BUILDS))
    assert best["pair_lanes"] in tuning.PAIR_LANES and best["kspace_split"] in tuning.KSPACE_SPLITS
    assert best["rho_width"] in tuning.RHO_WIDTHS and best["nlist_build"] in tuning.NLIST_BUILDS
    for s in systems:
        t = s.tuning()
        assert (t["pair_lanes"], t["kspace_split"], t["rho_width"]) == (
            best["pair_lanes"], best["kspace_split"], best["rho_width"])
        if best["nlist_build"]:
            assert t["nlist_build"] == best["nlist_build"]
    # fresh copies of the initial states with the chosen knobs
    for seed, f in zip((21, 22), f0):
        s = spce(nmol=300, L=20.8, cutoff=8.0, kmax=7, seed=seed)
        s.set_tuning(pair_lanes=best["pair_lanes"], kspace_split=best["kspace_split"], rho_width=best["rho_width"],
                     nlist_build=best["nlist_build"], rho_kind=best["rho_kind"])
        assert frel(s.forces(), f) < 1e-12
