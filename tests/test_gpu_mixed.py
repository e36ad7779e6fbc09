"""Mixed-precision form (FP32 pair/erfc/trig, FP64 positions, phase
reduction and accumulation) against the FP64 oracle.

Stated tolerances (the target is <= 1e-5 RMS relative force error on the
SPC/E configuration): force RMS relative error < 1e-5, virial relative
error (max entry / max |W|) < 1e-5, energy relative error < 1e-4, and an
NVE drift no more than twice the FP64 form's.  The energy bound is the
loosest: FP32 pair energies carry a systematic ~1e-7 relative rounding (of
q_i q_j, 1/(4 pi eps0), alpha) into sums whose terms are several times the
total in water, where real, self and k-space energies cancel (measured:
~4e-6 for 3000-molecule SPC/E, ~1e-7 for NaCl).
"""
import math

import numpy as np
import pytest

import lumol_oracle as lo
from helpers import nist_spce_system, random_molecular_system
from lumol_amd import MolecularDynamics, Simulation, VelocityVerlet
from lumol_amd._native import LG_PRECISION_FP64, LG_PRECISION_MIXED
from lumol_amd.workloads import argon, nacl, spce

pytestmark = pytest.mark.gpu

FORCE_RMS = 1e-5
ENERGY_REL = 1e-4
VIRIAL_REL = 1e-5


def errors(s):
    o = lo.OracleSystem(s, nthreads=16)
    s.set_precision(LG_PRECISION_MIXED)
    fg, fo = s.forces(), o.forces()
    d = fg - fo
    rms = math.sqrt((d * d).sum() / (fo * fo).sum())
    e = abs(s.potential_energy() - o.potential_energy()) / abs(o.potential_energy())
    wg, wo = s.virial(), o.virial()
    w = np.abs(wg - wo).max() / np.abs(wo).max()
    return rms, e, w


@pytest.mark.parametrize("builder", ["argon", "nacl_small", "nacl", "spce_small", "spce", "molecular_wolf"])
def test_mixed_error_bounds(builder):
    s = {"argon": lambda: argon(),
         "nacl_small": lambda: nacl(ncell=4, cutoff=10.0, kmax=7),
         "nacl": lambda: nacl(),
         "spce_small": lambda: spce(nmol=300, L=20.8, cutoff=8.0, kmax=7),
         "spce": lambda: spce(),
         "molecular_wolf": lambda: random_molecular_system(n_mol=40, L=16.0, coulomb="wolf")}[builder]()
    rms, e, w = errors(s)
    print(f"{builder}: force rms rel {rms:.2e}, energy rel {e:.2e}, virial rel {w:.2e}")
    assert rms < FORCE_RMS
    assert e < ENERGY_REL
    assert w < VIRIAL_REL


@pytest.mark.parametrize("name", ["spce-1.xyz", "spce-4.xyz"])
def test_mixed_nist_spce(name):
    rms, e, w = errors(nist_spce_system(name, 10.0))
    assert rms < FORCE_RMS and e < ENERGY_REL and w < VIRIAL_REL


def drift(form, nsteps=3000, every=100):
    """Slope of the total energy per atom (least squares), and its standard error."""
    s = nacl(ncell=4, cutoff=10.0, kmax=7)
    s.set_precision(form)
    sim = Simulation(MolecularDynamics(VelocityVerlet(1.0)))
    t, e = [], []
    for k in range(nsteps // every + 1):
        t.append(k * every)
        e.append(s.total_energy() / s.size())
        if k < nsteps // every:
            sim.run(s, every)
    t, e = np.array(t, float), np.array(e)
    A = np.vstack([t, np.ones_like(t)]).T
    coef, res, *_ = np.linalg.lstsq(A, e, rcond=None)
    resid = e - A @ coef
    se = math.sqrt((resid ** 2).sum() / (len(t) - 2) / ((t - t.mean()) ** 2).sum())
    return coef[0], se


def test_mixed_nve_drift_within_twice_fp64():
    d64, se64 = drift(LG_PRECISION_FP64)
    dmx, semx = drift(LG_PRECISION_MIXED)
    print(f"drift per atom-step: fp64 {d64:.3e} (+-{se64:.1e}), mixed {dmx:.3e} (+-{semx:.1e})")
    assert abs(dmx) <= 2 * abs(d64) + 3 * (se64 + semx)


@pytest.mark.parametrize("offset", [0.0, 700.0])
def test_mixed_infinite_cell_far_from_origin(offset):
    # An infinite cell: the mixed pair kernel reads FP32 coordinates relative
    # to the first atom's position (formed in FP64), so a charged LJ cluster
    # far from the coordinate origin keeps the error bound of one near it
    # (FP32 coordinates of magnitude 700 A alone would carry ~3e-5 A of
    # rounding per coordinate).
    from lumol_amd import LennardJones, PairInteraction, System, UnitCell, Wolf, units
    rng = np.random.default_rng(17)
    g = np.arange(6) * 3.6
    pos = np.array([(x, y, z) for x in g for y in g for z in g]) + rng.uniform(-0.3, 0.3, (216, 3))
    pos += offset * np.array([1.0, -0.6, 0.8])
    names = ["Na" if k % 2 == 0 else "Cl" for k in range(216)]
    charges = [0.7 if k % 2 == 0 else -0.7 for k in range(216)]
    s = System.from_arrays(names, pos, UnitCell.infinite(), charges=charges)
    lj = PairInteraction(LennardJones(3.3, units.from_(0.5, "kJ/mol")), 9.0)
    for pair in (("Na", "Na"), ("Na", "Cl"), ("Cl", "Cl")):
        s.set_pair_potential(pair, lj)
    s.set_coulomb_potential(Wolf(9.0))
    # forces and energy (the reference has no virial for an infinite cell,
    # compute.rs:198)
    o = lo.OracleSystem(s, nthreads=16)
    s.set_precision(LG_PRECISION_MIXED)
    fg, fo = s.forces(), o.forces()
    d = fg - fo
    rms = math.sqrt((d * d).sum() / (fo * fo).sum())
    e = abs(s.potential_energy() - o.potential_energy()) / abs(o.potential_energy())
    print(f"infinite cell, offset {offset}: force rms rel {rms:.2e}, energy rel {e:.2e}")
    assert rms < FORCE_RMS
    assert e < ENERGY_REL
