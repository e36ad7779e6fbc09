"""The host-side propagators (lumol_amd/propagators.py) pinned on the CPU:
they step around force, energy and virial evaluations, which here come from
the oracle through a stand-in with the System's interface, so that the
host arithmetic is checked against what the mathematics fixes:

* Verlet: the Stormer recurrence on a harmonic spring, time reversibility;
* Rescale: the reference's known answers (lumol-sim/tests/thermostats.rs:34-50);
* CSVR: the canonical distribution of the kinetic energy (mean dof kT/2,
  variance dof (kT)^2 / 2) on an ideal gas (thermostats.rs:79-117 recipe);
* Berendsen barostat: eta^3 = 1 - beta_T / tau (P0 - P) from the pressure
  of the drifted state, positions scaled by eta and the cell by eta^3;
* Rotate: Matrix3::rotation's doc value (types/matrix.rs:136-147),
  orthonormality, rigid rotation about the centre of mass;
* Resize: the cost beta (dE + P dV) - N ln(V'/V) against energies evaluated
  directly, centres of mass scaled with the cell, molecules rigid;
* Rewrap: controls.rs:123-131;
* the host Monte Carlo loop: a Null potential accepts every move, and the
  amplitude update moves the acceptance toward its target.
"""
import math

import numpy as np
import pytest

import lumol_oracle as lo
from lumol_amd import (Harmonic, LennardJones, MonteCarlo, NullPotential, PairInteraction, Simulation, System,
                       Translate, UnitCell, units)
from lumol_amd import propagators as P
from lumol_amd.consts import K_BOLTZMANN


class OracleBacked:
    """The evaluation interface of lumol_amd.System, answered by the oracle:
    positions and velocities live here, forces/energies/virials come from
    the oracle's restatement of the reference."""

    def __init__(self, system, nthreads=4):
        self.spec = system
        self.names = list(system.names)
        self._x = np.array(system.positions, dtype=float)
        self._v = np.array(system.velocities, dtype=float)
        self.masses = np.array(system.masses, dtype=float)
        self.cell = system.cell
        self.o = lo.OracleSystem(system, nthreads=nthreads)
        self.simulated_degrees_of_freedom = "particles"
        self.external_temperature = None

    def size(self):
        return len(self.masses)

    @property
    def positions(self):
        return self._x.copy()

    @positions.setter
    def positions(self, x):
        self._x = np.array(x, dtype=float).reshape(-1, 3)
        self.o.set_positions(self._x)

    @property
    def velocities(self):
        return self._v.copy()

    @velocities.setter
    def velocities(self, v):
        self._v = np.array(v, dtype=float).reshape(-1, 3)
        self.o.set_velocities(self._v)

    def cell_changed(self):
        self.o.set_cell(*self.cell.lengths)

    def molecules(self):
        return self.spec.molecules()

    def maximum_cutoff(self):
        return self.spec.maximum_cutoff()

    def forces(self):
        return self.o.forces()

    def potential_energy(self):
        return self.o.potential_energy()

    def kinetic_energy(self):
        return 0.5 * float((self.masses[:, None] * self._v * self._v).sum())

    def degrees_of_freedom(self):
        return 3 * self.size()

    def temperature(self):
        return 2.0 * self.kinetic_energy() / (self.degrees_of_freedom() * K_BOLTZMANN)

    def volume(self):
        a, b, c = self.cell.lengths
        return a * b * c

    def virial(self):
        return self.o.virial()

    def pressure(self):
        return (self.degrees_of_freedom() * K_BOLTZMANN * self.temperature() + np.trace(self.virial())) / (
            3.0 * self.volume())

    def stress(self):
        kin = (self.masses[:, None, None] * self._v[:, :, None] * self._v[:, None, :]).sum(0)
        return (kin + self.virial()) / self.volume()

    def cache_init(self):
        self.o.cache_init()
        return self.o.cache_energy()

    def cache_energy(self):
        return self.o.cache_energy()

    def move_cost(self, molid, newpos):
        return self.o.move_cost(molid, newpos)

    def move_commit(self, wrap=True):
        self.o.move_commit(wrap)
        self._x = self.o.positions()

    def move_discard(self):
        self.o.move_discard()


def ideal_gas(n_side=6, L=20.0, temperature=300.0, seed=1):
    """Helium atoms on a grid with Null interactions (thermostats.rs:12-30)."""
    from lumol_amd.input import boltzmann_velocities
    pos = [(i * 2.0, j * 2.0, k * 2.0) for i in range(n_side) for j in range(n_side) for k in range(n_side)]
    s = System.from_arrays(["He"] * len(pos), pos, UnitCell.cubic(L))
    s.set_pair_potential(("He", "He"), PairInteraction(NullPotential(), 3.0))
    boltzmann_velocities(s, temperature, seed=seed)
    return s


# --- Verlet --------------------------------------------------------------------------------------------
def spring(x0=1.5, k=0.3, r=1.2):
    s = System.from_arrays(["A", "A"], [(0.0, 0.0, 0.0), (r, 0.0, 0.0)], UnitCell.cubic(20.0), masses=[2.0, 2.0])
    s.set_pair_potential(("A", "A"), PairInteraction(Harmonic(k=k, x0=x0), 9.0))
    return s


def test_verlet_is_the_stormer_recurrence():
    k, x0, dt = 0.3, 1.5, 0.7
    b = OracleBacked(spring(x0=x0, k=k))
    b.velocities = np.array([[0.01, 0.0, 0.0], [-0.01, 0.0, 0.0]])
    integ = P.Verlet(dt)
    integ.setup(b)
    # separation d: d'' = -(2k/m)(d - x0); the recurrence on d itself
    d_prev = 1.2 - (-0.02) * dt
    d = 1.2
    for _ in range(50):
        integ.integrate(b)
        d_new = 2 * d - d_prev + dt * dt * (-(2 * k / 2.0) * (d - x0))
        d_prev, d = d, d_new
        got = b.positions[1, 0] - b.positions[0, 0]
        assert abs(got - d) <= 1e-12 * abs(d), (got, d)
    v_expected = (d - d_prev) / (2 * dt)  # not the last step's (central difference of the last pair)
    assert math.isfinite(v_expected)


def test_verlet_time_reversible():
    b = OracleBacked(spring())
    b.velocities = np.array([[0.02, 0.01, 0.0], [-0.02, -0.01, 0.0]])
    x0 = b.positions
    integ = P.Verlet(0.5)
    integ.setup(b)
    for _ in range(40):
        integ.integrate(b)
    # reverse: swap the stored previous positions with the current ones
    integ.prevpos, cur = b.positions, integ.prevpos
    b.positions = cur
    for _ in range(40):
        integ.integrate(b)
    assert np.abs(integ.prevpos - x0).max() < 1e-10


# --- thermostats ----------------------------------------------------------------------------------
def test_rescale_known_answers():
    # lumol-sim/tests/thermostats.rs:34-50
    b = OracleBacked(ideal_gas())
    assert abs(b.temperature() - 300.0) < 1e-9
    P.RescaleThermostat(250.0, 100.0).apply(b)
    assert abs(b.temperature() - 300.0) < 1e-12
    P.RescaleThermostat(250.0, 10.0).apply(b)
    assert abs(b.temperature() - 250.0) < 1e-9
    # default tolerance: 5% of the target
    t = P.RescaleThermostat(260.0)
    assert t.tolerance == 0.05 * 260.0


def test_csvr_samples_the_canonical_kinetic_energy():
    # ideal gas: no forces, so only the thermostat changes K; its stationary
    # distribution is Gamma(dof/2, kT): mean dof kT/2, variance dof (kT)^2/2
    b = OracleBacked(ideal_gas(n_side=4))
    T = 250.0
    th = P.CSVRThermostat(T, 10.0, seed=3)
    for _ in range(200):
        th.apply(b)
    ks = []
    for _ in range(4000):
        th.apply(b)
        ks.append(b.kinetic_energy())
    ks = np.array(ks)
    dof = b.degrees_of_freedom()
    kT = K_BOLTZMANN * T
    mean, var = ks.mean(), ks.var()
    # correlated samples (tau = 10 steps): ~4000 / 20 independent ones
    neff = len(ks) / 20.0
    sd_mean = math.sqrt(dof * kT * kT / 2.0 / neff)
    assert abs(mean - dof * kT / 2.0) < 5 * sd_mean, (mean, dof * kT / 2)
    assert abs(var / (dof * kT * kT / 2.0) - 1.0) < 0.25, var


def test_berendsen_thermostat_factor():
    # thermostats.rs:113-121: lambda = sqrt(1 + (T0 / T - 1) / tau)
    b = OracleBacked(ideal_gas())
    v0 = b.velocities
    P.HostBerendsen(250.0, 10.0).apply(b)
    lam = math.sqrt(1.0 + (250.0 / 300.0 - 1.0) / 10.0)
    assert np.abs(b.velocities - lam * v0).max() <= 1e-15 * np.abs(v0).max()


# --- Berendsen barostat ----------------------------------------------------------------------------
def lj_fluid(n_side=5, L=11.0, seed=2):
    rng = np.random.default_rng(seed)
    pos = [((i + 0.5) * L / n_side, (j + 0.5) * L / n_side, (k + 0.5) * L / n_side)
           for i in range(n_side) for j in range(n_side) for k in range(n_side)]
    pos = np.array(pos) + rng.uniform(-0.2, 0.2, size=(len(pos), 3))
    s = System.from_arrays(["Ar"] * len(pos), pos, UnitCell.cubic(L))
    s.set_pair_potential(("Ar", "Ar"), PairInteraction(LennardJones(3.4, units.from_(1.0, "kJ/mol")), 5.0))
    s.velocities = rng.normal(scale=2e-3, size=(len(pos), 3))
    return s


def test_berendsen_barostat_step():
    b = OracleBacked(lj_fluid())
    P0 = units.from_(500.0, "bar")
    baro = P.BerendsenBarostat(1.0, P0, 1000.0)
    baro.setup(b)
    L0 = b.cell.lengths[0]
    x0, v0 = b.positions, b.velocities
    baro.integrate(b)  # eta = 1: cell unchanged, plain velocity-Verlet drift
    assert b.cell.lengths[0] == L0
    # the pressure that set eta: drifted positions, half-kicked velocities
    ref = OracleBacked(lj_fluid())
    ref.positions = x0 + v0 * 1.0
    ref.velocities = v0
    eta3 = 1.0 - P.WATER_COMPRESSIBILITY / 1000.0 * (P0 - ref.pressure())
    assert abs(baro.eta - float(np.cbrt(eta3))) <= 1e-15
    eta = baro.eta
    x1, v1, a1 = b.positions, b.velocities, baro.acc
    baro.integrate(b)
    # second step: positions scaled by eta, the cell lengths by eta^3
    assert abs(b.cell.lengths[0] - L0 * eta ** 3) <= 1e-13 * L0
    vh = v1 + 0.5 * 1.0 * a1
    assert np.abs(b.positions - (x1 * eta + vh * 1.0)).max() <= 1e-12 * np.abs(x1).max()


# --- Monte Carlo moves ------------------------------------------------------------------------------
def test_rotation_matrix_doc_value():
    # types/matrix.rs:136-147: x axis (length 2), 90 deg, applied to e3 -> e2
    r = P.rotation_matrix([2.0, 0.0, 0.0], math.radians(90.0))
    v = np.round((r @ np.array([0.0, 0.0, 1.0])) * 1e8) / 1e8
    assert v.tolist() == [0.0, 1.0, 0.0]
    rng = np.random.default_rng(0)
    for _ in range(5):
        m = P.rotation_matrix(rng.normal(size=3), rng.uniform(-3, 3))
        assert np.abs(m @ m.T - np.eye(3)).max() < 1e-14
        assert abs(np.linalg.det(m) - 1.0) < 1e-14


def water_like(n_mol=8, L=14.0):
    from helpers import random_molecular_system
    return random_molecular_system(n_mol=n_mol, L=L, coulomb="wolf")


def test_rotate_is_rigid_about_the_com():
    s = water_like()
    b = OracleBacked(s)
    b.cache_init()
    mv = P.Rotate(0.8)
    rng = np.random.default_rng(4)
    mols = s.molecules()
    for _ in range(6):
        mv.prepare(b, rng, mols)
        a, e = mols[mv.molid]
        old = b.positions[a:e]
        m = b.masses[a:e]
        com_old = (m[:, None] * old).sum(0) / m.sum()
        com_new = (m[:, None] * mv.newpos).sum(0) / m.sum()
        assert np.abs(com_new - com_old).max() < 1e-12
        dold = np.linalg.norm(old[:, None] - old[None], axis=-1)
        dnew = np.linalg.norm(mv.newpos[:, None] - mv.newpos[None], axis=-1)
        assert np.abs(dold - dnew).max() < 1e-12
        # the cost is the energy change of the rotated configuration
        c = mv.cost(b, 1.0)
        e0 = b.potential_energy()
        x = b.positions
        x[a:e] = mv.newpos
        ref = OracleBacked(s)
        ref.positions = x
        assert abs(c - (ref.potential_energy() - e0)) <= 1e-9 * abs(e0)
        mv.restore(b)


def test_resize_cost_and_rigid_molecules():
    s = water_like()
    b = OracleBacked(s)
    b.cache_init()
    pressure = units.from_(200.0, "bar")
    mv = P.Resize(pressure, 30.0)
    rng = np.random.default_rng(6)
    mols = s.molecules()
    x0 = b.positions
    e0 = b.cache_energy()
    V0 = b.volume()
    mv.prepare(b, rng, mols)
    beta = 1.0 / (K_BOLTZMANN * 300.0)
    c = mv.cost(b, beta)
    V1 = b.volume()
    f = b.cell.lengths[0] / s.cell.lengths[0]
    assert abs(f ** 3 * V0 - V1) <= 1e-12 * V0
    x1 = b.positions
    for a, e in mols:
        m = b.masses[a:e]
        c0 = (m[:, None] * x0[a:e]).sum(0) / m.sum()
        c1 = (m[:, None] * x1[a:e]).sum(0) / m.sum()
        assert np.abs(c1 - c0 * f).max() < 1e-12 * max(1.0, np.abs(c0).max())
        assert np.abs((x1[a:e] - c1) - (x0[a:e] - c0)).max() < 1e-12
    ref = OracleBacked(s)
    ref.cell = b.cell
    ref.cell_changed()
    ref.positions = x1
    e1 = ref.potential_energy()
    expected = beta * (e1 - e0 + pressure * (V1 - V0)) - len(mols) * math.log(V1 / V0)
    assert abs(c - expected) <= 1e-9 * max(1.0, abs(expected))
    mv.restore(b)
    assert b.volume() == V0 and np.array_equal(b.positions, x0)


def test_rewrap_known_answer():
    # controls.rs:123-131
    s = System.from_arrays(["Ag", "Ag"], [(0.0, 0.0, 0.0), (15.0, 0.0, 0.0)], UnitCell.cubic(10.0))
    b = OracleBacked(s)
    P.Rewrap().control(b)
    assert b.positions.tolist() == [[0.0, 0.0, 0.0], [5.0, 0.0, 0.0]]


def test_host_monte_carlo_loop():
    # a Null potential accepts every move (cost 0); Rotate + Translate mixed
    s = ideal_gas(n_side=3)
    b = OracleBacked(s)
    mc = MonteCarlo(300.0, seed=9)
    mc.add(Translate(1.0), 1.0)
    mc.add(P.Rotate(0.5), 1.0)
    sim = Simulation(mc)
    b.cache_init()
    sim._run_mc_host(b, 200)
    assert mc.attempts == 200 and mc.accepted == 200
    # amplitude adaptation toward a target acceptance on an LJ fluid
    s2 = lj_fluid(n_side=3, L=11.0)
    b2 = OracleBacked(s2)
    mc2 = MonteCarlo(120.0, seed=2)
    mc2.add(Translate(3.0), 1.0, 0.5)
    mc2.set_amplitude_update_frequency(50)
    sim2 = Simulation(mc2)
    b2.cache_init()
    sim2._run_mc_host(b2, 600)
    delta = sim2._host[0][0].move.delta
    assert delta < 3.0  # a 3 A amplitude in a dense fluid accepts far below 50%
