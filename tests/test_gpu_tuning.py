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
    kind = 1 + (lanes + split) % 2
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
    assert best["kspace_split"] in tuning.KSPACE_SPLITS and best["rho_kind"] in (1, 2)
    assert best["rho_width"] in tuning.RHO_WIDTHS if best["rho_kind"] == 1 else best["rho_width"] == 0
    assert frel(s.forces(), f0) < 1e-12
    # a second system of the same bucket reuses the choice
    t = spce(nmol=300, L=20.8, cutoff=8.0, kmax=7, seed=9)
    assert tuning.autotune(t) == best
    assert tuning.bucket(t) in tuning.cached()
