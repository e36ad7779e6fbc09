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
import math
import os

import numpy as np
import pytest

from helpers import GOLDEN
from lumol_amd.consts import K_BOLTZMANN
from lumol_amd.input import Input
from test_gpu_parity import check_system as _check_system

pytestmark = pytest.mark.gpu

SCEN = os.path.join(GOLDEN, "scenarios")


def check_system(s, nthreads=16):
    """Parity of the initial configuration at 1e-10.  Some scenarios start on
    a perfect lattice (md-nacl) or as a dilute gas with no pair inside the
    cutoff (mc-helium), where every force is rounding noise: forces are then
    held to 1e-10 of the force scale |E| / N per angstrom (or 1e-300)."""
    from lumol_oracle import OracleSystem
    e = abs(OracleSystem(s, nthreads=4).potential_energy())
    return _check_system(s, nthreads=nthreads, force_floor=max(e / s.size(), 1e-300))

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
    check_system(s)
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
    check_system(s)
    config.simulation.run(s, config.nsteps)
    pv = s.pressure() * s.volume()
    nkt = s.size() * K_BOLTZMANN * 300.0
    assert abs(pv - nkt) / pv < 2e-2
    assert config.simulation.propagator.accepted > 0


# --- integrators, barostats and moves stepped on the host around device evaluations ---------------
def _collect(config, start):
    """Run the simulation step by step, collecting pressure and temperature
    after step `start` as the reference's Collector output does
    (tests/utils/mod.rs:14-54; System.step counts from 1 after a step)."""
    s, sim = config.system, config.simulation
    pressures, temperatures = [], []
    for step in range(1, config.nsteps + 1):
        sim.run(s, 1)
        if step >= start:
            pressures.append(s.pressure())
            temperatures.append(s.temperature())
    return np.mean(pressures), np.mean(temperatures)


def test_md_helium_verlet_constant_energy():
    # tests/md-helium.rs:33-48
    config = Input(os.path.join(SCEN, "md-helium", "nve-verlet.toml")).read()
    s = config.system
    e_initial = s.total_energy()
    config.simulation.run(s, config.nsteps)
    drift = abs((e_initial - s.total_energy()) / s.total_energy())
    print(f"md-helium/nve-verlet: |dE/E| = {drift:.2e}", flush=True)
    assert drift < 1e-2


def test_md_helium_berendsen_barostat():
    # tests/md-helium.rs:86-107: 5000 bar within 5%, 273 K within 1%
    from lumol_amd import units
    config = Input(os.path.join(SCEN, "md-helium", "npt-berendsen-barostat.toml")).read()
    p, t = _collect(config, 4000)
    print(f"md-helium/npt: <P> = {units.to(p, 'bar'):.1f} bar, <T> = {t:.2f} K", flush=True)
    assert abs(p - units.from_(5000.0, "bar")) / units.from_(5000.0, "bar") < 5e-2
    assert abs(t - 273.0) / 273.0 < 1e-2


def test_md_nacl_anisotropic_berendsen():
    # tests/md-nacl.rs:33-57: 50000 bar within 2e-3, 273 K within 1%
    from lumol_amd import units
    config = Input(os.path.join(SCEN, "md-nacl", "npt-wolf-small.toml")).read()
    p, t = _collect(config, 9000)
    print(f"md-nacl/npt-wolf-small: <P> = {units.to(p, 'bar'):.1f} bar, <T> = {t:.2f} K", flush=True)
    assert abs(p - units.from_(50000.0, "bar")) / units.from_(50000.0, "bar") < 2e-3
    assert abs(t - 273.0) / 273.0 < 1e-2


def _collect_series(config, start):
    s, sim = config.system, config.simulation
    out = []
    for step in range(1, config.nsteps + 1):
        sim.run(s, 1)
        if step >= start:
            out.append(s.pressure())
    return np.array(out)


def test_mc_argon_npt():
    # tests/mc-argon.rs:13-32 checks the mean pressure after step 500 at 1%
    # of 200 bar, on the reference's one random stream.  This run draws
    # other numbers, so the bar is the larger of 1% and three standard
    # errors of the mean (ten blocks), stated with the result.
    from lumol_amd import units
    config = Input(os.path.join(SCEN, "mc-argon", "npt.toml")).read()
    check_system(config.system)
    p = _collect_series(config, 500)
    blocks = np.array([b.mean() for b in np.array_split(p, 10)])
    sem = blocks.std(ddof=1) / math.sqrt(len(blocks))
    target = units.from_(200.0, "bar")
    bar = max(1e-2 * target, 3 * sem)
    print(f"mc-argon/npt: <P> = {units.to(p.mean(), 'bar'):.2f} bar, 3 sem = {units.to(3 * sem, 'bar'):.2f} bar",
          flush=True)
    assert abs(p.mean() - target) < bar


@pytest.mark.parametrize("name", ["nvt-wolf.toml", "nvt-ewald.toml"])
def test_mc_water_nvt(name):
    # tests/mc-water.rs:13-24, 48-58: the runs complete (the reference checks nothing more)
    config = Input(os.path.join(SCEN, "mc-water", name)).read()
    s = config.system
    check_system(s)
    config.simulation.run(s, config.nsteps)
    mc = config.simulation.propagator
    assert 0 < mc.accepted < mc.attempts == config.nsteps
    assert np.isfinite(s.potential_energy())
    # the running cache matches a fresh evaluation
    assert abs(s.cache_energy() - s.potential_energy()) <= 1e-8 * abs(s.potential_energy())


@pytest.mark.parametrize("name", ["npt-wolf.toml", "npt-ewald.toml"])
def test_mc_water_npt(name):
    # tests/mc-water.rs:26-46, 60-80: 1000 bar within 800 bar over the last 5000 steps
    from lumol_amd import units
    config = Input(os.path.join(SCEN, "mc-water", name)).read()
    p, _ = _collect(config, config.nsteps - 5000)
    print(f"mc-water/{name}: <P> = {units.to(p, 'bar'):.1f} bar", flush=True)
    assert abs(p - units.from_(1000.0, "bar")) < units.from_(800.0, "bar")
