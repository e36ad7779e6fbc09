"""MD chunks replayed from a captured HIP graph give the same bits as the
same kernels launched one by one: several replays of one graph, a second
graph for a shorter final chunk, eager work (force and energy evaluations,
position reads) between the runs, both precision forms, VelocityVerlet and
LeapFrog with a Berendsen thermostat."""
import numpy as np
import pytest

from lumol_amd import BerendsenThermostat, LeapFrog, MolecularDynamics, Simulation, VelocityVerlet
from lumol_amd._native import LG_PRECISION_FP64, LG_PRECISION_MIXED
from lumol_amd.workloads import nacl, spce

pytestmark = pytest.mark.gpu


def water(precision):
    s = spce(nmol=216, L=18.64, cutoff=8.0, kmax=6)
    s.set_precision(precision)
    return s, Simulation(MolecularDynamics(VelocityVerlet(0.5)))


def salt(precision):
    s = nacl(ncell=4, cutoff=10.0, kmax=7)
    s.set_precision(precision)
    prop = MolecularDynamics(LeapFrog(1.0))
    prop.set_thermostat(BerendsenThermostat(1200.0, 100.0))
    return s, Simulation(prop)


@pytest.mark.parametrize("make", [water, salt])
@pytest.mark.parametrize("precision", [LG_PRECISION_FP64, LG_PRECISION_MIXED])
def test_graph_replay_matches_eager(make, precision):
    runs = []
    for graph in (1, 2):  # 1: every kernel launched, 2: chunks replayed from a graph
        s, sim = make(precision)
        s.set_tuning(md_graph=graph)
        trace = []
        for steps in (450, 200, 130):  # chunks of 200, 200, 50 | 200 | 130
            sim.run(s, steps)
            trace.append((s.positions.copy(), s.velocities.copy(), s.forces().copy(), s.potential_energy()))
        assert s.tuning()["md_graph"] == graph
        runs.append(trace)
    for a, b in zip(*runs):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
