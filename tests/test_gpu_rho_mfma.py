"""rho(k) on the matrix cores (a complex product over atoms on v_mfma
16x16x4) against the row form and the oracle: k-space energy, virial and
forces at the FP64 parity bar, the mixed form within its force-error bar,
across kmax (column tiles from one to many, rows ikx up to 15), atom counts
below one block's stage and off the stage width, and the fall-back to the row
form above kmax 16.  The row form stages the phase rows in LDS up to kmax 12
(63 rows) and reads them from the table above, so kmax 3..12 and 16..17
cover both of its loops."""
import numpy as np
import pytest

import lumol_oracle as lo
from lumol_amd._native import LG_PRECISION_MIXED
from lumol_amd.workloads import nacl, spce

pytestmark = pytest.mark.gpu

CASES = {
    # (maker, kind in effect when the matrix-core form is asked for)
    "nacl-64-kmax3": (lambda: nacl(ncell=2, cutoff=5.0, kmax=3), 2),
    "nacl-512-kmax7": (lambda: nacl(ncell=4, cutoff=10.0, kmax=7), 2),
    "nacl-4096-kmax9": (lambda: nacl(ncell=8, cutoff=12.0, kmax=9), 2),
    # 63 phase rows: the largest table the row form stages in LDS
    "nacl-512-kmax12": (lambda: nacl(ncell=4, cutoff=10.0, kmax=12), 2),
    "water-1500-kmax10": (lambda: spce(nmol=500, L=24.7, cutoff=8.0, kmax=10), 2),
    "water-3000-kmax16": (lambda: spce(nmol=1000, L=31.07, cutoff=7.0, kmax=16), 2),
    "water-ragged-kmax5": (lambda: spce(nmol=333, L=21.6, cutoff=7.0, kmax=5), 2),
}


def frel(f, g):
    return np.abs(f - g).max() / np.abs(g).max()


def run(make, kind, mixed=False):
    s = make()
    if mixed:
        s.set_precision(LG_PRECISION_MIXED)
    s.set_tuning(rho_kind=kind)
    f = s.forces()
    e = s.energy_terms()
    w = s.virial()
    return s, f, e, w


@pytest.mark.parametrize("case", list(CASES))
def test_matrix_core_rho_fp64(case):
    make, kind = CASES[case]
    s2, f2, e2, w2 = run(make, 2)
    assert s2.tuning()["rho_kind"] == kind
    _, f1, e1, w1 = run(make, 1)
    # same sum, other order
    assert frel(f2, f1) < 1e-12
    assert abs(e2.coulomb_kspace - e1.coulomb_kspace) <= 1e-12 * abs(e1.coulomb_kspace)
    assert np.abs(w2 - w1).max() <= 1e-12 * np.abs(w1).max()
    o = lo.OracleSystem(s2, nthreads=8)
    fo, eo, wo = o.forces(), o.energy_terms(), o.virial()
    assert frel(f2, fo) < 1e-10
    assert abs(e2.total - eo["total"]) <= 1e-10 * abs(eo["total"])
    assert np.abs(w2 - wo).max() <= 1e-10 * np.abs(wo).max()


@pytest.mark.parametrize("case", ["nacl-4096-kmax9", "water-1500-kmax10", "water-3000-kmax16"])
def test_matrix_core_rho_mixed(case):
    make, _ = CASES[case]
    s2, f2, e2, _ = run(make, 2, mixed=True)
    assert s2.tuning()["rho_kind"] == 2
    _, f1, e1, _ = run(make, 1, mixed=True)
    fo = lo.OracleSystem(s2, nthreads=8).forces()
    rms = lambda a, b: np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum())
    assert rms(f2, fo) < 1e-5
    # no worse than the row form by more than FP32 rounding
    assert rms(f2, fo) < 2.0 * rms(f1, fo) + 1e-6
    assert abs(e2.coulomb_kspace - e1.coulomb_kspace) <= 1e-5 * abs(e1.coulomb_kspace)


@pytest.mark.parametrize("case", ["nacl-512-kmax7", "water-1500-kmax10", "nacl-512-kmax12"])
@pytest.mark.parametrize("mixed", [False, True])
def test_staged_rows_equal_table_rows(case, mixed):
    # the row form with its phase rows staged in LDS (rho_kind 1) and read
    # from the table (rho_kind 3) adds the same terms in the same order:
    # the same bits, in both forms
    make, _ = CASES[case]
    s1, f1, e1, w1 = run(make, 1, mixed=mixed)
    s3, f3, e3, w3 = run(make, 3, mixed=mixed)
    assert s3.tuning()["rho_kind"] == 3 and s1.tuning()["rho_kind"] == 1
    assert np.array_equal(f1, f3)
    assert e1.coulomb_kspace == e3.coulomb_kspace
    assert np.array_equal(w1, w3)


def test_kmax_above_16_uses_rows():
    s = nacl(ncell=4, cutoff=10.0, kmax=17)
    s.set_tuning(rho_kind=2)
    assert s.tuning()["rho_kind"] == 1
    fo = lo.OracleSystem(s, nthreads=8).forces()
    assert frel(s.forces(), fo) < 1e-10


def test_default_is_rows():
    assert spce(nmol=500, L=24.7, cutoff=8.0, kmax=10).tuning()["rho_kind"] == 1


def test_matrix_core_rho_md_matches_rows():
    # 40 VV steps of water: the two forms stay within summation-order noise
    from lumol_amd import MolecularDynamics, Simulation, VelocityVerlet
    out = []
    for kind in (1, 2):
        s = spce(nmol=500, L=24.7, cutoff=8.0, kmax=8)
        s.set_tuning(rho_kind=kind)
        Simulation(MolecularDynamics(VelocityVerlet(0.5))).run(s, 40)
        out.append(s.positions.copy())
    assert np.abs(out[0] - out[1]).max() < 1e-9
