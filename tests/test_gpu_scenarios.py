"""The reference's own simulation scenarios (tests/golden/scenarios, copied
from its tests/data), read by lumol_amd.input and run on the device with the
reference's integration-test bars:

  md-nacl   nve-wolf-small 1e-4, nve-ewald-small 5e-3, nve-ewald-kspace 5e-3   tests/md-nacl.rs:17-94
  md-water  nve-ewald 1e-1, nve-wolf 3e-2                                      tests/md-water.rs:13-43
  md-helium velocity-verlet 5e-3, leap-frog 5e-3, shifted 2e-3, table 5e-3;
            perfect gas |pV - NkT| < 1e-3                                      tests/md-helium.rs:17-150
  md-butane 1e-3, md-methane 1e-2                                              tests/md-butane.rs, md-methane.rs
  mc-helium perfect gas |pV - NkT| / pV < 2e-2                                 tests/mc-helium.rs:16-35

(relative change of the total energy over the run).  Before each run, the
initial configuration's forces, energies and virial are held to the
oracle at 1e-10.  Velocities come from the input's `init` temperature with
this package's seeded generator, not the reference's stream.
"""
import os

import numpy as np
import pytest

from helpers import GOLDEN
from lumol_amd.consts import K_BOLTZMANN
from lumol_amd.input import Input
from test_gpu_parity import check_system

pytestmark = pytest.mark.gpu

SCEN = os.path.join(GOLDEN, "scenarios")

DRIFT = [
    ("md-nacl", "nve-wolf-small.toml", 1e-4),
    ("md-nacl", "nve-ewald-small.toml", 5e-3),
    ("md-nacl", "nve-ewald-kspace.toml", 5e-3),
    ("md-water", "nve-ewald.toml", 1e-1),
    ("md-water", "nve-wolf.toml", 3e-2),
    ("md-helium", "nve-velocity-verlet.toml", 5e-3),
    ("md-helium", "nve-leap-frog.toml", 5e-3),
    ("md-helium", "nve-shifted.toml", 2e-3),
    ("md-helium", "nve-table.toml", 5e-3),
    ("md-butane", "nve.toml", 1e-3),
    ("md-methane", "nve.toml", 1e-2),
]


@pytest.mark.parametrize("folder,name,bar", DRIFT, ids=[f"{a}/{b}" for a, b, _ in DRIFT])
def test_constant_energy(folder, name, bar):
    config = Input(os.path.join(SCEN, folder, name)).read()
    s = config.system
    check_system(s, nthreads=16)
    e_initial = s.total_energy()
    config.simulation.run(s, config.nsteps)
    e_final = s.total_energy()
    drift = abs((e_initial - e_final) / e_final)
    print(f"{folder}/{name}: {config.nsteps} steps, |dE/E| = {drift:.2e} (bar {bar:.0e})", flush=True)
    assert np.isfinite(e_final)
    assert drift < bar


def test_md_helium_perfect_gas():
    # tests/md-helium.rs:66-84
    config = Input(os.path.join(SCEN, "md-helium", "nve-perfect-gas.toml")).read()
    s = config.system
    config.simulation.run(s, config.nsteps)
    pv = s.pressure() * s.volume()
    nkt = s.size() * K_BOLTZMANN * s.temperature()
    assert abs(pv - nkt) < 1e-3


def test_mc_helium_perfect_gas():
    # tests/mc-helium.rs:16-35: pressure at the Monte Carlo temperature
    config = Input(os.path.join(SCEN, "mc-helium", "nvt.toml")).read()
    s = config.system
    check_system(s, nthreads=16)
    config.simulation.run(s, config.nsteps)
    pv = s.pressure() * s.volume()
    nkt = s.size() * K_BOLTZMANN * 300.0
    assert abs(pv - nkt) / pv < 2e-2
    assert config.simulation.propagator.accepted > 0
