"""Without a thermostat, lg_md_run gives the second half of a step and the
first half of the next one launch (k_md_fused).  The fused steps give the
same bits as steps launched one by one: runs of one step (no fusion inside
a run) against one run of many steps, VelocityVerlet and LeapFrog, both
precision forms, with and without k-space and intra-molecular terms."""
import numpy as np
import pytest

from lumol_amd import LeapFrog, MolecularDynamics, Simulation, VelocityVerlet
from lumol_amd._native import LG_PRECISION_FP64, LG_PRECISION_MIXED
from lumol_amd.workloads import lj_fluid, spce

pytestmark = pytest.mark.gpu


def water(precision, integrator):
    s = spce(nmol=216, L=18.64, cutoff=8.0, kmax=6)
    s.set_precision(precision)
    return s, Simulation(MolecularDynamics(integrator(0.5)))


def argon(precision, integrator):
    s = lj_fluid(2000)
    s.set_precision(precision)
    return s, Simulation(MolecularDynamics(integrator(2.0)))


@pytest.mark.parametrize("make", [water, argon])
@pytest.mark.parametrize("integrator", [VelocityVerlet, LeapFrog])
@pytest.mark.parametrize("precision", [LG_PRECISION_FP64, LG_PRECISION_MIXED])
def test_fused_steps_match_single_steps(make, integrator, precision):
    out = []
    for chunks in ([1] * 30, [30]):
        s, sim = make(precision, integrator)
        for c in chunks:
            sim.run(s, c)
        out.append((s.positions.copy(), s.velocities.copy(), s.potential_energy()))
    a, b = out
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert a[2] == b[2]
