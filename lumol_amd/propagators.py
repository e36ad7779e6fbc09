"""Propagator parts that run on the host around the device evaluations: the
integrators, thermostats, controls and Monte Carlo moves of lumol-sim that
the fused device MD step and the device Monte Carlo chain do not cover.

  Verlet, BerendsenBarostat, AnisoBerendsenBarostat   lumol-sim/src/md/integrators.rs:72-342
  VelocityVerlet, LeapFrog (host forms)               lumol-sim/src/md/integrators.rs:22-70, 125-169
  RescaleThermostat, BerendsenThermostat, CSVR        lumol-sim/src/md/thermostats.rs:40-217
  RemoveTranslation, RemoveRotation, Rewrap           lumol-sim/src/md/controls.rs:26-90
  Translate, Rotate, Resize moves, amplitude updates  lumol-sim/src/mc/moves/*.rs, mc/monte_carlo.rs:17-345

Every force, energy, virial and pressure these steps need comes from the
device (System.forces / potential_energy / pressure ...); the host does the
O(N) position and velocity updates, the cell scaling and the random draws.
Random numbers come from numpy's generator (seeded): the reference's
xorshift streams are not reproduced, only their distributions.
"""
import math

import numpy as np

from .consts import K_BOLTZMANN

# integrators.rs:171-173 (from DL_POLY)
WATER_COMPRESSIBILITY = 7372.0


# --- integrators -------------------------------------------------------------------------------
class HostVelocityVerlet:
    """VelocityVerlet (integrators.rs:22-70), stepping on the host with device
    forces; accelerations start at zero (:40-42)."""

    def __init__(self, timestep):
        self.timestep = float(timestep)
        self.acc = None

    def setup(self, system):
        self.acc = np.zeros((system.size(), 3))

    def integrate(self, system):
        dt = self.timestep
        m = system.masses[:, None]
        v = system.velocities + 0.5 * dt * self.acc
        system.velocities = v
        system.positions = system.positions + v * dt
        self.acc = system.forces() / m
        system.velocities = v + 0.5 * dt * self.acc


class HostLeapFrog:
    """LeapFrog (integrators.rs:125-169) on the host with device forces."""

    def __init__(self, timestep):
        self.timestep = float(timestep)
        self.acc = None

    def setup(self, system):
        self.acc = np.zeros((system.size(), 3))

    def integrate(self, system):
        dt = self.timestep
        dt2 = dt * dt
        system.positions = system.positions + system.velocities * dt + 0.5 * self.acc * dt2
        new = system.forces() / system.masses[:, None]
        system.velocities = system.velocities + 0.5 * (self.acc + new) * dt
        self.acc = new


class Verlet:
    """Position Verlet (integrators.rs:72-123): x(t - dt) = x - v dt at setup,
    then x' = 2x - x_prev + dt^2 F/m and v = (x' - x_prev) / (2 dt)."""

    def __init__(self, timestep):
        self.timestep = float(timestep)
        self.prevpos = None

    def setup(self, system):
        self.prevpos = system.positions - system.velocities * self.timestep

    def integrate(self, system):
        dt = self.timestep
        dt2 = dt * dt
        f = system.forces()
        x = system.positions
        new = 2.0 * x - self.prevpos + (dt2 / system.masses)[:, None] * f
        system.positions = new
        system.velocities = (new - self.prevpos) / (2.0 * dt)
        self.prevpos = x


def _check_cell(system, what):
    cut = system.maximum_cutoff()
    if cut is not None and any(0.5 * d <= cut for d in system.cell.lengths):
        raise ValueError(f"Tried to decrease the cell size in {what} but the new size is smaller than "
                         "the interactions cut off radius. You can try to increase the cell size or the "
                         "number of particles.")


class BerendsenBarostat:
    """Isotropic Berendsen barostat on velocity-Verlet (integrators.rs:171-250):
    eta^3 = 1 - beta_T / tau (P0 - P(t)) with the pressure of the drifted
    state; as the reference has it, positions are scaled by eta and the cell
    matrix by eta^3 (integrators.rs:213-220)."""

    def __init__(self, timestep, pressure, tau):
        self.timestep = float(timestep)
        self.pressure = float(pressure)
        self.tau = float(tau)
        self.acc = None
        self.eta = 1.0

    def setup(self, system):
        self.acc = np.zeros((system.size(), 3))

    def integrate(self, system):
        from .system import UnitCell
        dt = self.timestep
        v = system.velocities + 0.5 * dt * self.acc
        x = system.positions * self.eta + v * dt
        e3 = self.eta * self.eta * self.eta
        system.cell = UnitCell.ortho(*(L * e3 for L in system.cell.lengths))
        system.cell_changed()
        _check_cell(system, "Berendesen barostat")
        system.positions = x
        system.velocities = v
        eta3 = 1.0 - WATER_COMPRESSIBILITY / self.tau * (self.pressure - system.pressure())
        self.eta = float(np.cbrt(eta3))
        self.acc = system.forces() / system.masses[:, None]
        system.velocities = v + 0.5 * dt * self.acc


class AnisoBerendsenBarostat:
    """Anisotropic Berendsen barostat on velocity-Verlet
    (integrators.rs:252-342): eta = 1 - dt beta_T / tau (S0 - S(t)),
    symmetrised.  The device evaluates orthorhombic cells only, so the
    diagonal of eta scales the cell and the positions here; the off-diagonal
    entries, which would shear the cell, are left out (for a system whose
    stress is diagonal the two agree)."""

    def __init__(self, timestep, stress, tau):
        self.timestep = float(timestep)
        self.stress = np.array(stress, dtype=float).reshape(3, 3)
        self.tau = float(tau)
        self.acc = None
        self.eta = np.eye(3)

    @classmethod
    def hydrostatic(cls, timestep, pressure, tau):
        return cls(timestep, pressure * np.eye(3), tau)

    def setup(self, system):
        self.acc = np.zeros((system.size(), 3))

    def integrate(self, system):
        from .system import UnitCell
        dt = self.timestep
        d = np.diag(self.eta).copy()
        v = system.velocities + 0.5 * dt * self.acc
        x = system.positions * d[None, :] + v * dt
        system.cell = UnitCell.ortho(*(L * f for L, f in zip(system.cell.lengths, d)))
        system.cell_changed()
        _check_cell(system, "anisotropic Berendesen barostat")
        system.positions = x
        system.velocities = v
        factor = self.timestep * WATER_COMPRESSIBILITY / self.tau
        eta = np.eye(3) - factor * (self.stress - system.stress())
        for i in range(3):
            for j in range(i):
                eta[i][j] = 0.5 * (eta[i][j] + eta[j][i])
                eta[j][i] = eta[i][j]
        self.eta = eta
        self.acc = system.forces() / system.masses[:, None]
        system.velocities = v + 0.5 * dt * self.acc


# --- thermostats ----------------------------------------------------------------------------------
def scale_velocities(system, temperature):
    """velocities::scale (velocities.rs:14-21)."""
    factor = math.sqrt(temperature / system.temperature())
    system.velocities = system.velocities * factor


class RescaleThermostat:
    """Rescale to T when |T(t) - T| exceeds the tolerance, 5% of T by default
    (thermostats.rs:40-80)."""

    def __init__(self, temperature, tolerance=None):
        if temperature < 0:
            raise ValueError("The temperature must be positive in thermostats.")
        self.temperature = float(temperature)
        self.tolerance = 0.05 * self.temperature if tolerance is None else float(tolerance)

    def apply(self, system):
        if abs(system.temperature() - self.temperature) > self.tolerance:
            scale_velocities(system, self.temperature)


class HostBerendsen:
    """BerendsenThermostat::apply (thermostats.rs:113-121) on the host."""

    def __init__(self, temperature, tau):
        self.temperature = float(temperature)
        self.tau = float(tau)

    def apply(self, system):
        factor = math.sqrt(1.0 + (self.temperature / system.temperature() - 1.0) / self.tau)
        system.velocities = system.velocities * factor


class CSVRThermostat:
    """Canonical sampling through velocity rescaling (Bussi, Donadio and
    Parrinello 2007), thermostats.rs:123-217: one rescaling per step with the
    sum of dof - 1 squared normals drawn as 2 Gamma((dof - 1)/2) (plus one
    squared normal when dof - 1 is odd)."""

    def __init__(self, temperature, tau, seed=42):
        if temperature < 0:
            raise ValueError("The temperature must be positive in thermostats.")
        if tau < 1.0:
            raise ValueError("The timestep must be larger than 1 in CSVR thermostat.")
        self.target_kinetic_per_dof = K_BOLTZMANN * temperature / 2.0
        self.tau = float(tau)
        self.rng = np.random.default_rng(seed)

    def _sum_noises(self, n):
        if n == 0:
            return 0.0, 0.0
        rr = self.rng.standard_normal()
        if n == 1:
            return rr, rr * rr
        if n % 2 == 0:
            return rr, 2.0 * self.rng.gamma(n // 2, 1.0)
        return rr, 2.0 * self.rng.gamma((n - 1) // 2, 1.0) + rr * rr

    def apply(self, system):
        kinetic = system.kinetic_energy()
        kinetic_factor = self.target_kinetic_per_dof / kinetic
        exp_1 = math.exp(-1.0 / self.tau)
        exp_2 = (1.0 - exp_1) * kinetic_factor
        gauss, wiener = self._sum_noises(system.degrees_of_freedom() - 1)
        scale = exp_1 + exp_2 * (gauss * gauss + wiener) + 2.0 * gauss * math.sqrt(exp_1 * exp_2)
        system.velocities = system.velocities * math.sqrt(scale)


# --- controls ----------------------------------------------------------------------------------------
class _Every:
    def __init__(self, every=1):
        if int(every) < 1:
            raise ValueError("'every' must be at least 1")
        self.every = int(every)
        self.count = 0

    def control(self, system):
        # lumol-input/src/alternator.rs: the wrapped control runs every n-th step
        self.count += 1
        if self.count % self.every == 0:
            self._control(system)


class RemoveTranslation(_Every):
    """controls.rs:26-43"""

    def _control(self, system):
        from .input import remove_translation
        system.velocities = remove_translation(system.masses, system.velocities)


class RemoveRotation(_Every):
    """controls.rs:45-76"""

    def _control(self, system):
        from .input import remove_rotation
        system.velocities = remove_rotation(system.masses, system.positions, system.velocities)


class Rewrap(_Every):
    """Molecules' centres of mass back into the cell (controls.rs:78-90,
    molecules.rs:287-296)."""

    def _control(self, system):
        if system.cell.is_infinite():
            return
        x = system.positions.copy()
        m = system.masses
        L = np.array(system.cell.lengths)
        for a, b in system.molecules():
            total = 0.0
            com = np.zeros(3)
            for i in range(a, b):
                total += m[i]
                com += m[i] * x[i]
            com = com / total
            wrapped = com - np.floor(com / L) * L
            x[a:b] += wrapped - com
        system.positions = x


# --- Monte Carlo moves ----------------------------------------------------------------------------
def rotation_matrix(axis, angle):
    """Matrix3::rotation (types/matrix.rs:149-168)."""
    n = np.asarray(axis, dtype=float)
    n = n / math.sqrt(float(n @ n))
    s, c = math.sin(angle), math.cos(angle)
    omc = 1.0 - c
    xy, xz, yz = n[0] * n[1] * omc, n[0] * n[2] * omc, n[1] * n[2] * omc
    return np.array([[(n[0] * n[0]) * omc + c, xy + n[2] * s, xz - n[1] * s],
                     [xy - n[2] * s, (n[1] * n[1]) * omc + c, yz + n[0] * s],
                     [xz + n[1] * s, yz - n[0] * s, (n[2] * n[2]) * omc + c]])


def _com(x, m):
    total = 0.0
    com = np.zeros(3)
    for xi, mi in zip(x, m):
        total += mi
        com += mi * xi
    return com / total


class Rotate:
    """Rotation of a random molecule about its centre of mass by U(-theta,
    theta) around a uniform random axis (mc/moves/rotate.rs:20-125); no
    wrap on accept."""

    def __init__(self, theta):
        if not theta > 0:
            raise ValueError("theta must be positive in Rotate move")
        self.theta = float(theta)

    def prepare(self, system, rng, mols):
        self.molid = int(rng.integers(len(mols)))
        a, b = mols[self.molid]
        axis = rng.standard_normal(3)
        while float(axis @ axis) == 0.0:
            axis = rng.standard_normal(3)
        angle = rng.uniform(-self.theta, self.theta)
        x = system.positions[a:b]
        com = _com(x, system.masses[a:b])
        rot = rotation_matrix(axis, angle)
        self.newpos = np.array([com + rot @ (p - com) for p in x])
        return True

    def cost(self, system, beta):
        return beta * system.move_cost(self.molid, self.newpos)

    def apply(self, system):
        system.move_commit(False)

    def restore(self, system):
        system.move_discard()

    def update_amplitude(self, s, system):
        if s is not None:
            if math.degrees(abs(s * self.theta)) <= 180.0:
                self.theta *= s


class HostTranslate:
    """Translate (mc/moves/translate.rs:22-148) driven from the host:
    displacement U(-delta/sqrt(3), delta/sqrt(3)) per axis, centre of mass
    wrapped on accept."""

    def __init__(self, delta):
        if not delta > 0:
            raise ValueError("delta must be positive in Translate move")
        self.delta = float(delta)

    def prepare(self, system, rng, mols):
        self.molid = int(rng.integers(len(mols)))
        a, b = mols[self.molid]
        d = self.delta / math.sqrt(3.0)
        shift = rng.uniform(-d, d, size=3)
        self.newpos = system.positions[a:b] + shift
        return True

    def cost(self, system, beta):
        return beta * system.move_cost(self.molid, self.newpos)

    def apply(self, system):
        system.move_commit(True)

    def restore(self, system):
        system.move_discard()

    def update_amplitude(self, s, system):
        if s is not None:
            cut = system.maximum_cutoff()
            if cut is not None and self.delta * s > cut:
                return
            self.delta *= s


class Resize:
    """Volume change of U(-delta, delta) with the molecules' centres of mass
    scaled and the molecules moved rigidly (mc/moves/resize.rs:14-128); the
    cost is beta (dE + P dV) - N_mol ln(V'/V) with dE from the device
    (EnergyCache::move_all_molecules_cost, cache.rs:238-292: intra-molecular
    terms do not change)."""

    def __init__(self, pressure, delta):
        if not delta > 0:
            raise ValueError("delta must be positive in Resize move")
        self.pressure = float(pressure)
        self.delta = float(delta)

    def prepare(self, system, rng, mols):
        from .system import UnitCell
        if system.cell.is_infinite():
            raise ValueError("cannot use `Resize` move with infinite simulation cell")
        dv = rng.uniform(-self.delta, self.delta)
        self.old_cell = system.cell
        self.old_pos = system.positions.copy()
        self.old_energy = system.cache_energy()
        volume = system.volume()
        f = float(np.cbrt((volume + dv) / volume))
        system.cell = UnitCell.ortho(*(L * f for L in self.old_cell.lengths))
        system.cell_changed()
        _check_cell(system, "Monte Carlo Resize move")
        x = self.old_pos.copy()
        m = system.masses
        old_L = np.array(self.old_cell.lengths)
        new_L = np.array(system.cell.lengths)
        for a, b in mols:
            com = _com(x[a:b], m[a:b])
            x[a:b] += (com / old_L) * new_L - com
        system.positions = x
        return True

    def cost(self, system, beta):
        new_energy = system.cache_init()
        self.new_energy = new_energy
        nmol = len(system.molecules())
        dv = system.volume() - self.old_cell.volume()
        return beta * (new_energy - self.old_energy + self.pressure * dv) - nmol * math.log(
            system.volume() / self.old_cell.volume())

    def apply(self, system):
        pass

    def restore(self, system):
        system.cell = self.old_cell
        system.cell_changed()
        system.positions = self.old_pos
        system.cache_init()

    def update_amplitude(self, s, system):
        if s is not None:
            self.delta *= s


class MoveCounter:
    """Move of monte_carlo.rs:17-105: counts and the amplitude scaling factor."""

    def __init__(self, move, target_acceptance=None):
        if target_acceptance is not None and not 0.0 < target_acceptance < 1.0:
            raise ValueError(f"target acceptance ratio must be between 0 and 1, got {target_acceptance}")
        self.move = move
        self.target = target_acceptance
        self.total_attempted = self.total_accepted = self.attempted = self.accepted = 0

    def accept(self):
        self.total_attempted += 1
        self.attempted += 1
        self.accepted += 1
        self.total_accepted += 1

    def reject(self):
        self.total_attempted += 1
        self.attempted += 1

    def scaling_factor(self):
        if self.target is None or self.attempted == 0:
            return None
        q = self.accepted / self.attempted / self.target
        return 1.2 if q > 1.2 else 0.8 if q < 0.8 else q

    def acceptance(self):
        return 0.0 if self.total_attempted == 0 else self.total_accepted / self.total_attempted
