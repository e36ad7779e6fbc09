"""Host-side mirror of lumol-core's system and potential API, backed by the
gfx950 library.

Names and argument meaning follow the reference so that code written
against lumol reads the same (paths relative to the reference checkout):

  LennardJones / Harmonic / NullPotential,
  Buckingham / BornMayerHuggins / Morse /
  Gaussian / Mie / CosineHarmonic / Torsion lumol-core/src/energy/functions.rs
  TableComputation                          lumol-core/src/energy/computations.rs
  PairInteraction (new, shifted, tail)      lumol-core/src/energy/pairs.rs:57-297
  PairRestriction                           lumol-core/src/energy/restrictions.rs:13-33
  Ewald / SharedEwald / Wolf                lumol-core/src/energy/global/{ewald,wolf}.rs
  UnitCell (cubic, ortho, infinite)         lumol-core/src/sys/config/cells.rs
  System (add_bond, set_*_potential,
          forces, potential_energy, virial) lumol-core/src/sys/system.rs
  VelocityVerlet / LeapFrog                 lumol-sim/src/md/integrators.rs
  BerendsenThermostat                       lumol-sim/src/md/thermostats.rs:91-121
  MolecularDynamics / Simulation            lumol-sim/src/md/molecular_dynamics.rs,
                                            lumol-sim/src/simulations.rs

Every evaluation runs on the GPU through the C ABI; invalid input raises
ValueError (where the reference panics) or LumolGpuError.
"""
import ctypes
import math

import numpy as np

from . import _native as nat

from . import unitconv as _units
from .consts import FOUR_PI_EPSILON_0, K_BOLTZMANN  # noqa: F401

# Standard atomic weights (IUPAC) of the elements, the values of lumol's mass
# table (lumol-core/src/sys/config/mass.rs); a particle whose name is not an
# element gets mass 0 (particles.rs:61).
ATOMIC_MASSES = {
    "H": 1.008, "He": 4.002602, "Li": 6.94, "Be": 9.012182, "B": 10.81, "C": 12.011,
    "N": 14.007, "O": 15.999, "F": 18.9984032, "Ne": 20.1797, "Na": 22.98976928, "Mg": 24.305,
    "Al": 26.9815386, "Si": 28.085, "P": 30.973762, "S": 32.06, "Cl": 35.45, "Ar": 39.948,
    "K": 39.0983, "Ca": 40.078, "Sc": 44.955912, "Ti": 47.867, "V": 50.9415, "Cr": 51.9961,
    "Mn": 54.938045, "Fe": 55.845, "Co": 58.933195, "Ni": 58.6934, "Cu": 63.546, "Zn": 65.38,
    "Ga": 69.723, "Ge": 72.63, "As": 74.9216, "Se": 78.96, "Br": 79.904, "Kr": 83.798,
    "Rb": 85.4678, "Sr": 87.62, "Y": 88.90585, "Zr": 91.224, "Nb": 92.90638, "Mo": 95.96,
    "Tc": 97.0, "Ru": 101.07, "Rh": 102.9055, "Pd": 106.42, "Ag": 107.8682, "Cd": 112.411,
    "In": 114.818, "Sn": 118.71, "Sb": 121.76, "Te": 127.6, "I": 126.90447, "Xe": 131.293,
    "Cs": 132.9054519, "Ba": 137.327, "La": 138.90547, "Ce": 140.116, "Pr": 140.90765, "Nd": 144.242,
    "Pm": 145.0, "Sm": 150.36, "Eu": 151.964, "Gd": 157.25, "Tb": 158.92535, "Dy": 162.5,
    "Ho": 164.93032, "Er": 167.259, "Tm": 168.93421, "Yb": 173.054, "Lu": 174.9668, "Hf": 178.49,
    "Ta": 180.94788, "W": 183.84, "Re": 186.207, "Os": 190.23, "Ir": 192.217, "Pt": 195.084,
    "Au": 196.966569, "Hg": 200.592, "Tl": 204.38, "Pb": 207.2, "Bi": 208.9804, "Po": 209.0,
    "At": 210.0, "Rn": 222.0, "Fr": 223.0, "Ra": 226.0, "Ac": 227.0, "Th": 232.03806,
    "Pa": 231.03588, "U": 238.02891, "Np": 237.0, "Pu": 244.0, "Am": 243.0, "Cm": 247.0,
    "Bk": 247.0, "Cf": 251.0, "Es": 252.0, "Fm": 257.0, "Md": 258.0, "No": 259.0,
    "Lr": 262.0, "Rf": 267.0, "Db": 270.0, "Sg": 271.0, "Bh": 270.0, "Hs": 277.0,
    "Mt": 276.0, "Ds": 281.0, "Rg": 282.0, "Cn": 285.0, "Uut": 285.0, "Fl": 289.0,
    "Mc": 289.0, "Lv": 293.0, "Ts": 294.0, "Og": 294.0,
}


class units:
    """Conversion to and from lumol's internal units (lumol-core/src/units.rs);
    unit strings such as "kcal/mol/A^2" or "(kJ/mol/A)^2"."""
    from_ = staticmethod(_units.from_)
    from_str = staticmethod(_units.from_str)
    to = staticmethod(_units.to)


# ---------------------------------------------------------------------------------
# potentials
# ---------------------------------------------------------------------------------
class NullPotential:
    """V = 0 (functions.rs:30-50)."""
    kind = nat.LG_POT_NULL

    def params(self):
        return (0.0, 0.0)


class LennardJones:
    """V = 4 eps ((s/r)^12 - (s/r)^6) (functions.rs:52-107)."""
    kind = nat.LG_POT_LJ

    def __init__(self, sigma, epsilon):
        self.sigma = float(sigma)
        self.epsilon = float(epsilon)

    def params(self):
        return (self.sigma, self.epsilon)


class Harmonic:
    """V = k/2 (x - x0)^2 (functions.rs:109-160)."""
    kind = nat.LG_POT_HARMONIC

    def __init__(self, k, x0):
        self.k = float(k)
        self.x0 = float(x0)

    def params(self):
        return (self.k, self.x0)


class Buckingham:
    """V = A exp(-r/rho) - C/r^6 (functions.rs:260-327; its tail virial keeps
    the reference's constant `+ 8.0`)."""
    kind = nat.LG_POT_BUCKINGHAM

    def __init__(self, a, c, rho):
        self.a, self.c, self.rho = float(a), float(c), float(rho)

    def params(self):
        return (self.a, self.c, self.rho)


class BornMayerHuggins:
    """V = A exp((sigma - r)/rho) - C/r^6 + D/r^8 (functions.rs:330-395)."""
    kind = nat.LG_POT_BORN

    def __init__(self, a, c, d, sigma, rho):
        self.a, self.c, self.d, self.sigma, self.rho = float(a), float(c), float(d), float(sigma), float(rho)

    def params(self):
        return (self.a, self.c, self.d, self.sigma, self.rho)


class Morse:
    """V = depth (1 - exp(a (x0 - x)))^2 (functions.rs:397-437)."""
    kind = nat.LG_POT_MORSE

    def __init__(self, a, x0, depth):
        self.a, self.x0, self.depth = float(a), float(x0), float(depth)

    def params(self):
        return (self.a, self.x0, self.depth)


class Gaussian:
    """V = -a exp(-b x^2), b > 0 (functions.rs:439-499)."""
    kind = nat.LG_POT_GAUSSIAN

    def __init__(self, a, b):
        if not b > 0:
            raise ValueError('"b" has to be positive in Gaussian potential')
        self.a, self.b = float(a), float(b)

    def params(self):
        return (self.a, self.b)


class Mie:
    """V = eps n/(n-m) (n/m)^(m/(n-m)) ((s/r)^n - (s/r)^m), m < n
    (functions.rs:501-592)."""
    kind = nat.LG_POT_MIE

    def __init__(self, sigma, epsilon, n, m):
        if not m < n:
            raise ValueError("The repulsive exponent n has to be larger than the attractive exponent m")
        self.sigma, self.epsilon, self.n, self.m = float(sigma), float(epsilon), float(n), float(m)

    def params(self):
        return (self.sigma, self.epsilon, self.n, self.m)


class CosineHarmonic:
    """V = k/2 (cos x - cos x0)^2, angles and dihedrals (functions.rs:162-210)."""
    kind = nat.LG_POT_COSINE_HARMONIC

    def __init__(self, k, x0):
        self.k, self.x0 = float(k), float(x0)

    def params(self):
        return (self.k, self.x0)


class Torsion:
    """V = k (1 + cos(n x - delta)), dihedrals (functions.rs:212-258)."""
    kind = nat.LG_POT_TORSION

    def __init__(self, k, delta, n):
        if int(n) != n or n < 0:
            raise ValueError("the multiplicity of a torsion must be a non-negative integer")
        self.k, self.delta, self.n = float(k), float(delta), int(n)

    def params(self):
        return (self.k, self.delta, float(self.n))


class TableComputation:
    """A pair potential evaluated from tables of `size` points on [0, max)
    with linear interpolation, zero from the last bin on; tails stay those
    of the potential (computations.rs:70-175)."""

    def __init__(self, potential, size, max):
        if int(size) < 1:
            raise ValueError("the table needs at least one point")
        if not max > 0:
            raise ValueError("the table maximum must be positive")
        self.potential = potential
        self.table = (int(size), float(max))
        self.kind = potential.kind

    def params(self):
        return self.potential.params()


class PairRestriction:
    """Restriction codes of restrictions.rs:13-33 (Scale14 carries a factor)."""
    NONE = 0
    INTRA_MOLECULAR = 1
    INTER_MOLECULAR = 2
    EXCLUDE12 = 3
    EXCLUDE13 = 4
    EXCLUDE14 = 5

    def __init__(self, code, scaling=1.0):
        self.code = code
        self.scaling = scaling

    @classmethod
    def scale14(cls, scaling):
        if not 0.0 <= scaling <= 1.0:
            raise ValueError("scaling must be in [0, 1]")
        return cls(6, scaling)


def _restriction(r):
    if isinstance(r, PairRestriction):
        return r.code, r.scaling
    return int(r), 1.0


class PairInteraction:
    """Potential + cutoff (+ shift, tail, restriction) (pairs.rs:27-297)."""

    def __init__(self, potential, cutoff):
        self.potential = potential
        self._cutoff = float(cutoff)
        self.shifted_ = False
        self.tail = False
        self.restriction_ = PairRestriction.NONE
        self.scale14 = 1.0

    @classmethod
    def shifted(cls, potential, cutoff):
        p = cls(potential, cutoff)
        p.shifted_ = True
        return p

    def enable_tail_corrections(self):
        self.tail = True

    def set_restriction(self, restriction):
        self.restriction_, self.scale14 = _restriction(restriction)

    def restriction(self):
        return self.restriction_

    def cutoff(self):
        return self._cutoff


class Ewald:
    """Ewald summation (ewald.rs:228-305).  alpha None selects pi / cutoff."""

    def __init__(self, cutoff, kmax, alpha=None):
        if cutoff < 0:
            raise ValueError("the cutoff can not be negative in Ewald")
        if alpha is not None and alpha < 0:
            raise ValueError("alpha can not be negative in Ewald")
        if kmax == 0:
            raise ValueError("kmax can not be 0 in Ewald")
        self.cutoff = float(cutoff)
        self.kmax = int(kmax)
        self.alpha = math.pi / cutoff if alpha is None else float(alpha)
        self.restriction = PairRestriction.NONE

    def set_restriction(self, restriction):
        code, _ = _restriction(restriction)
        if code == 6:
            raise ValueError("Scaling restriction scheme using Ewald are not implemented")
        self.restriction = code


def SharedEwald(ewald):
    """The reference's thread-safe wrapper (ewald.rs:847); the GPU handle is
    already serialised, so the wrapped solver is used as is."""
    return ewald


class Wolf:
    """Wolf summation (wolf.rs:53-84)."""

    def __init__(self, cutoff):
        if not cutoff > 0:
            raise ValueError("Got a negative cutoff in Wolf summation")
        self.cutoff = float(cutoff)
        self.restriction = PairRestriction.NONE
        self.scale14 = 1.0

    def set_restriction(self, restriction):
        self.restriction, self.scale14 = _restriction(restriction)


class UnitCell:
    """Orthorhombic or infinite cell (cells.rs:31-120)."""

    def __init__(self, lengths):
        self.lengths = tuple(float(x) for x in lengths)

    @classmethod
    def cubic(cls, length):
        if not length > 0:
            raise ValueError("Cell lengths must be positive")
        return cls((length, length, length))

    @classmethod
    def ortho(cls, a, b, c):
        if not (a > 0 and b > 0 and c > 0):
            raise ValueError("Cell lengths must be positive")
        return cls((a, b, c))

    @classmethod
    def infinite(cls):
        return cls((0.0, 0.0, 0.0))

    def is_infinite(self):
        return self.lengths == (0.0, 0.0, 0.0)

    def volume(self):
        a, b, c = self.lengths
        return a * b * c


# ---------------------------------------------------------------------------------
# System
# ---------------------------------------------------------------------------------
class System:
    """Particles, topology and interactions; evaluation on the GPU.

    Positions and velocities are host arrays until the first evaluation; after
    that the device copy is authoritative and the properties read it back.
    """

    def __init__(self, cell=None, device=0):
        self.cell = cell if cell is not None else UnitCell.infinite()
        self.device = device
        self.stream_slot = 0  # library context (HIP stream) of the device
        self.names = []
        self._positions = np.zeros((0, 3))
        self._velocities = np.zeros((0, 3))
        self.charges = np.zeros(0)
        self.masses = np.zeros(0)
        self.bonds = []
        self.kinds = {}  # name -> kind index, in order of first use
        self.pairs = {}
        self.bond_potentials = {}
        self.angle_potentials = {}
        self.dihedral_potentials = {}
        self.coulomb = None
        # DegreesOfFreedom of system.rs:47-50, set by the propagator's setup
        self.simulated_degrees_of_freedom = "particles"
        # external temperature of system.rs:47 (set by a Monte Carlo setup)
        self.external_temperature = None
        self.precision = nat.LG_PRECISION_FP64
        self._handle = None
        self._dirty = True
        self._device_state = False

    # -- construction ----------------------------------------------------------
    def _kind(self, name):
        if name not in self.kinds:
            self.kinds[name] = len(self.kinds)
        return self.kinds[name]

    def add_particle(self, name, position=(0, 0, 0), velocity=(0, 0, 0), charge=0.0, mass=None):
        self._pull()
        self._kind(name)
        self.names.append(name)
        self._positions = np.vstack([self._positions, np.asarray(position, float).reshape(1, 3)])
        self._velocities = np.vstack([self._velocities, np.asarray(velocity, float).reshape(1, 3)])
        self.charges = np.append(self.charges, float(charge))
        self.masses = np.append(self.masses, ATOMIC_MASSES.get(name, 0.0) if mass is None else float(mass))
        self._dirty = True

    @classmethod
    def from_arrays(cls, names, positions, cell=None, charges=None, masses=None, velocities=None, device=0):
        s = cls(cell, device)
        for n in names:
            s._kind(n)
        s.names = list(names)
        s._positions = np.array(positions, dtype=float).reshape(-1, 3)
        n = len(s.names)
        s._velocities = (np.zeros((n, 3)) if velocities is None
                         else np.array(velocities, dtype=float).reshape(-1, 3))
        s.charges = np.zeros(n) if charges is None else np.array(charges, dtype=float)
        s.masses = (np.array([ATOMIC_MASSES.get(x, 0.0) for x in names]) if masses is None
                    else np.array(masses, dtype=float))
        return s

    def size(self):
        return len(self.names)

    def add_bond(self, i, j):
        if i == j or not (0 <= i < self.size() and 0 <= j < self.size()):
            raise ValueError("invalid bond")
        self.bonds.append((min(i, j), max(i, j)))
        self._dirty = True

    def set_pair_potential(self, names, interaction):
        a, b = names
        if not self.cell.is_infinite() and interaction.cutoff() > min(self.cell.lengths) / 2:
            raise ValueError("The cutoff for this pair potential is bigger than half the cell length")
        self.pairs[(a, b)] = interaction
        self._kind(a)
        self._kind(b)
        self._dirty = True

    def set_bond_potential(self, names, potential):
        self.bond_potentials[tuple(names)] = potential
        for n in names:
            self._kind(n)
        self._dirty = True

    def set_angle_potential(self, names, potential):
        self.angle_potentials[tuple(names)] = potential
        for n in names:
            self._kind(n)
        self._dirty = True

    def set_dihedral_potential(self, names, potential):
        self.dihedral_potentials[tuple(names)] = potential
        for n in names:
            self._kind(n)
        self._dirty = True

    def set_coulomb_potential(self, potential):
        # system.rs:157-167
        if (potential is not None and not self.cell.is_infinite()
                and potential.cutoff > min(self.cell.lengths) / 2):
            raise ValueError("The cutoff for the coulomb potential is bigger than half the cell length")
        self.coulomb = potential
        self._dirty = True

    def cell_changed(self):
        """Push a new `cell` (assigned to the attribute) to the device: the
        k-vectors, cell list and cutoff checks follow at the next evaluation."""
        if self._handle is not None and not self._dirty:
            a, b, c = self.cell.lengths
            nat.check(nat.lib().lg_system_set_cell(self._handle, a, b, c))
        else:
            self._dirty = True

    def maximum_cutoff(self):
        """Interactions::maximum_cutoff (interactions.rs:166-196): the largest
        pair or coulomb cutoff, None without potentials."""
        cuts = [p.cutoff() for p in self.pairs.values()]
        if self.coulomb is not None:
            cuts.append(self.coulomb.cutoff)
        return max(cuts) if cuts else None

    def set_precision(self, precision):
        self.precision = precision
        self._dirty = True

    def set_tuning(self, pair_lanes=0, kspace_split=0, rho_width=0, nlist_build=0, md_graph=0, rho_kind=0):
        """Launch knobs of the device kernels (0 = built-in choice); see
        lumol_amd.tuning.autotune.  rho_width: atoms per block of the row form
        of the rho(k) kernel; nlist_build: 1 lanes per atom, 2 one workgroup
        per cell over the cell list, 3 direct (workgroups per cell over the
        atoms' cell tags, up to 32768 atoms); md_graph: 1 eager launches (default), 2 HIP-graph replay of
        MD chunks; rho_kind: 1 row form (phase rows staged in LDS), 2
        matrix-core form (kmax <= 16), 3 row form without the staging."""
        self.tuning_ = (int(pair_lanes), int(kspace_split), int(rho_width), int(nlist_build), int(md_graph),
                        int(rho_kind))
        if self._handle is not None and not self._dirty:
            t = nat.Tuning(*self.tuning_)
            nat.check(nat.lib().lg_system_set_tuning(self._handle, ctypes.byref(t)))

    def tuning(self):
        """The launch knobs in effect: dict(pair_lanes, kspace_split, rho_width,
        nlist_build, md_graph, rho_kind)."""
        t = nat.Tuning()
        nat.check(nat.lib().lg_system_get_tuning(self._sync(), ctypes.byref(t)))
        return {"pair_lanes": t.pair_lanes, "kspace_split": t.kspace_split, "rho_width": t.rho_width,
                "nlist_build": t.nlist_build, "md_graph": t.md_graph, "rho_kind": t.rho_kind}

    # -- state -----------------------------------------------------------------
    @property
    def positions(self):
        self._pull()
        return self._positions

    @positions.setter
    def positions(self, value):
        self._pull()
        self._positions = np.array(value, dtype=float).reshape(-1, 3)
        if self._handle is not None:
            nat.check(nat.lib().lg_system_set_positions(self._handle, nat.dptr(self._positions)))

    @property
    def velocities(self):
        self._pull()
        return self._velocities

    @velocities.setter
    def velocities(self, value):
        self._pull()
        self._velocities = np.array(value, dtype=float).reshape(-1, 3)
        if self._handle is not None:
            nat.check(nat.lib().lg_system_set_velocities(self._handle, nat.dptr(self._velocities)))

    def _pull(self):
        if self._handle is not None and self._device_state:
            L = nat.lib()
            pos = np.zeros((self.size(), 3))
            vel = np.zeros((self.size(), 3))
            nat.check(L.lg_system_get_positions(self._handle, nat.dptr(pos)))
            nat.check(L.lg_system_get_velocities(self._handle, nat.dptr(vel)))
            self._positions, self._velocities = pos, vel
            self._device_state = False

    def particles_changed(self):
        """Push host-side edits of charges or masses to the device."""
        self._pull()
        self._dirty = True

    # -- device handle -------------------------------------------------------------
    def _sync(self):
        L = nat.lib()
        if self._handle is not None and not self._dirty:
            return self._handle
        self._pull()
        if self._handle is not None and self._handle_size != self.size():
            L.lg_system_destroy(self._handle)
            self._handle = None
        if self._handle is None:
            h = ctypes.c_void_p()
            nat.check(L.lg_system_create(nat.context(self.device, self.stream_slot), self.size(), ctypes.byref(h)))
            self._handle = h
            self._handle_size = self.size()
        h = self._handle
        a, b, c = self.cell.lengths
        nat.check(L.lg_system_set_cell(h, a, b, c))
        kinds = np.array([self.kinds[n] for n in self.names], dtype=np.uint32)
        nat.check(L.lg_system_set_kinds(h, nat.u32ptr(kinds), max(1, len(self.kinds))))
        q = np.ascontiguousarray(self.charges, dtype=float)
        nat.check(L.lg_system_set_charges(h, nat.dptr(q)))
        m = np.ascontiguousarray(self.masses, dtype=float)
        if np.all(m > 0):
            nat.check(L.lg_system_set_masses(h, nat.dptr(m)))
        pos = np.ascontiguousarray(self._positions)
        vel = np.ascontiguousarray(self._velocities)
        nat.check(L.lg_system_set_positions(h, nat.dptr(pos)))
        nat.check(L.lg_system_set_velocities(h, nat.dptr(vel)))
        bonds = np.array(self.bonds, dtype=np.uint32).reshape(-1, 2)
        nat.check(L.lg_system_set_bonds(h, len(bonds), nat.u32ptr(np.ascontiguousarray(bonds))))
        for (na, nb), inter in self.pairs.items():
            pot = inter.potential
            pr = np.array(pot.params() or (0.0,), dtype=float)
            table = getattr(pot, "table", None)
            if table is not None:
                nat.check(L.lg_pair_set_table(h, self.kinds[na], self.kinds[nb], pot.kind, nat.dptr(pr),
                                              table[0], table[1], inter.cutoff(), int(inter.shifted_),
                                              int(inter.tail), int(inter.restriction_), float(inter.scale14)))
            else:
                nat.check(L.lg_pair_set(h, self.kinds[na], self.kinds[nb], pot.kind, nat.dptr(pr),
                                        inter.cutoff(), int(inter.shifted_), int(inter.tail),
                                        int(inter.restriction_), float(inter.scale14)))
        for names, pot in self.bond_potentials.items():
            pr = np.array(pot.params() or (0.0,), dtype=float)
            nat.check(L.lg_bond_set(h, self.kinds[names[0]], self.kinds[names[1]], pot.kind, nat.dptr(pr)))
        for names, pot in self.angle_potentials.items():
            pr = np.array(pot.params() or (0.0,), dtype=float)
            k = [self.kinds[x] for x in names]
            nat.check(L.lg_angle_set(h, k[0], k[1], k[2], pot.kind, nat.dptr(pr)))
        for names, pot in self.dihedral_potentials.items():
            pr = np.array(pot.params() or (0.0,), dtype=float)
            k = [self.kinds[x] for x in names]
            nat.check(L.lg_dihedral_set(h, k[0], k[1], k[2], k[3], pot.kind, nat.dptr(pr)))
        if isinstance(self.coulomb, Ewald):
            nat.check(L.lg_coulomb_ewald(h, self.coulomb.cutoff, self.coulomb.kmax, self.coulomb.alpha,
                                         self.coulomb.restriction))
        elif isinstance(self.coulomb, Wolf):
            nat.check(L.lg_coulomb_wolf(h, self.coulomb.cutoff, self.coulomb.restriction, self.coulomb.scale14))
        else:
            nat.check(L.lg_coulomb_clear(h))
        nat.check(L.lg_system_set_precision(h, self.precision))
        t = nat.Tuning(*getattr(self, "tuning_", (0, 0, 0, 0, 0, 0)))
        nat.check(L.lg_system_set_tuning(h, ctypes.byref(t)))
        self._dirty = False
        self._device_state = True
        return h

    def close(self):
        if self._handle is not None:
            self._pull()
            nat.lib().lg_system_destroy(self._handle)
            self._handle = None
            self._dirty = True

    def __del__(self):
        try:
            if self._handle is not None and nat._lib is not None:
                nat._lib.lg_system_destroy(self._handle)
        except Exception:
            pass

    # -- evaluation (System::forces / potential_energy / virial) -------------------------
    def forces(self):
        h = self._sync()
        out = np.zeros((self.size(), 3))
        nat.check(nat.lib().lg_forces(h, nat.dptr(out), 0))
        return out

    def energy_terms(self):
        h = self._sync()
        t = nat.EnergyTerms()
        nat.check(nat.lib().lg_potential_energy(h, ctypes.byref(t)))
        return t

    def potential_energy(self):
        return self.energy_terms().total

    def atomic_virial(self):
        """AtomicVirial (compute.rs:195-255)."""
        h = self._sync()
        out = np.zeros(9)
        nat.check(nat.lib().lg_virial(h, nat.dptr(out)))
        return out.reshape(3, 3)

    def molecular_virial(self):
        """MolecularVirial (compute.rs:278-364)."""
        h = self._sync()
        out = np.zeros(9)
        nat.check(nat.lib().lg_molecular_virial(h, nat.dptr(out)))
        return out.reshape(3, 3)

    def virial(self):
        """Virial (compute.rs:372-381): the molecular definition when the
        simulated degrees of freedom are molecules (Monte Carlo moves of whole
        molecules), else the atomic one."""
        if self.simulated_degrees_of_freedom == "molecules":
            return self.molecular_virial()
        return self.atomic_virial()

    def degrees_of_freedom(self):
        """system.rs:249-255: 3 per particle, or 3 per molecule."""
        if self.simulated_degrees_of_freedom == "molecules":
            return 3 * len(self.molecules())
        return 3 * self.size()

    def kinetic_energy(self):
        h = self._sync()
        out = ctypes.c_double()
        nat.check(nat.lib().lg_kinetic_energy(h, ctypes.byref(out)))
        return out.value

    def total_energy(self):
        return self.kinetic_energy() + self.potential_energy()

    def temperature(self):
        """System::temperature (system.rs:273-278): the external temperature a
        Monte Carlo propagator sets, else the kinetic temperature."""
        if self.external_temperature is not None:
            return self.external_temperature
        h = self._sync()
        out = ctypes.c_double()
        nat.check(nat.lib().lg_temperature(h, ctypes.byref(out)))
        return out.value

    def volume(self):
        return self.cell.volume()

    def pressure_at_temperature(self, temperature):
        """PressureAtTemperature (compute.rs:396-405): (dof kB T + tr W) / (3 V)."""
        if self.cell.is_infinite():
            raise ValueError("Can not compute pressure for infinite cell")
        if not temperature >= 0:
            raise ValueError("the temperature must be positive")
        w = self.virial()
        return (self.degrees_of_freedom() * K_BOLTZMANN * temperature + np.trace(w)) / (3.0 * self.volume())

    def pressure(self):
        """System::pressure (system.rs:290-302): at the external temperature
        when one is set, else Pressure (compute.rs:418-425) at the kinetic
        temperature."""
        return self.pressure_at_temperature(self.temperature())

    def stress_at_temperature(self, temperature):
        """StressAtTemperature (compute.rs:442-452): (dof/3 kB T 1 + W) / V."""
        if self.cell.is_infinite():
            raise ValueError("Can not compute stress for infinite cell")
        if not temperature >= 0:
            raise ValueError("the temperature must be positive")
        w = self.virial()
        kinetic = self.degrees_of_freedom() / 3.0 * K_BOLTZMANN * temperature * np.eye(3)
        return (kinetic + w) / self.volume()

    def stress(self):
        """System::stress (system.rs:304-315): StressAtTemperature at the
        external temperature when one is set, else Stress (compute.rs:465-478),
        (sum_i m_i v_i (x) v_i + W) / V."""
        if self.external_temperature is not None:
            return self.stress_at_temperature(self.external_temperature)
        if self.cell.is_infinite():
            raise ValueError("Can not compute stress for infinite cell")
        v = self.velocities
        kinetic = np.zeros((3, 3))
        for m, vi in zip(self.masses, v):
            kinetic += m * np.outer(vi, vi)
        return (kinetic + self.virial()) / self.volume()

    def pairs_within(self, cutoff):
        """Inter-molecular pairs (i < j) with r < cutoff found by the device
        neighbour search, sorted; same distance code as the force kernels."""
        h = self._sync()
        count = ctypes.c_uint64()
        nat.check(nat.lib().lg_pairs_within(h, cutoff, None, 0, ctypes.byref(count)))
        out = np.zeros((count.value, 2), dtype=np.uint32)
        nat.check(nat.lib().lg_pairs_within(h, cutoff, nat.u32ptr(out), count.value, ctypes.byref(count)))
        order = np.lexsort((out[:, 1], out[:, 0]))
        return out[order]

    def pair_count_within(self, cutoff):
        """Number of inter-molecular pairs (i < j) with r < cutoff found by the
        device neighbour search."""
        count = ctypes.c_uint64()
        nat.check(nat.lib().lg_pairs_within(self._sync(), cutoff, None, 0, ctypes.byref(count)))
        return count.value

    def profile_begin(self, classes):
        """Time the given kernel classes (names of _native.KERNEL_CLASSES)
        with HIP events until profile_end()."""
        mask = 0
        for c in classes:
            mask |= 1 << nat.KERNEL_CLASSES.index(c)
        nat.check(nat.lib().lg_profile_begin(self._sync(), mask))

    def profile_end(self):
        p = nat.Profile()
        nat.check(nat.lib().lg_profile_end(self._sync(), ctypes.byref(p)))
        return {c: (p.ms[i], p.launches[i]) for i, c in enumerate(nat.KERNEL_CLASSES)}

    def ewald_kvector_count(self):
        h = self._sync()
        out = ctypes.c_int()
        nat.check(nat.lib().lg_ewald_kvector_count(h, ctypes.byref(out)))
        return out.value

    # -- Monte Carlo energy cache (lumol-core/src/sys/cache.rs) -------------------------
    def molecules(self):
        """(start, end) index range of every molecule: the connected
        components of the bond graph, each a contiguous range of atoms."""
        n = self.size()
        parent = list(range(n))

        def find(x):
            while parent[x] != x:
                parent[x] = parent[parent[x]]
                x = parent[x]
            return x

        for i, j in self.bonds:
            a, b = find(i), find(j)
            if a != b:
                parent[max(a, b)] = min(a, b)
        ranges = []
        for i in range(n):
            if find(i) == i:
                ranges.append([i, i + 1])
            else:
                ranges[-1][1] = i + 1
        return [tuple(r) for r in ranges]

    def cache_init(self):
        """EnergyCache::init (cache.rs:75-98); returns the cached energy."""
        e = ctypes.c_double()
        nat.check(nat.lib().lg_cache_init(self._sync(), ctypes.byref(e)))
        return e.value

    def cache_energy(self):
        """EnergyCache::energy (cache.rs:101-113)."""
        e = ctypes.c_double()
        nat.check(nat.lib().lg_cache_energy(self._sync(), ctypes.byref(e)))
        return e.value

    def move_cost(self, molid, newpos):
        """EnergyCache::move_molecule_cost (cache.rs:145-220): energy change
        of moving molecule `molid` to `newpos`; stages the move."""
        newpos = np.ascontiguousarray(newpos, dtype=float).reshape(-1, 3)
        c = ctypes.c_double()
        nat.check(nat.lib().lg_move_cost(self._sync(), int(molid), nat.dptr(newpos), ctypes.byref(c)))
        return c.value

    def move_commit(self, wrap=True):
        """EnergyCache::update (cache.rs:118-127) and the COM wrap of
        Translate::apply (translate.rs:115-127)."""
        nat.check(nat.lib().lg_move_commit(self._sync(), int(bool(wrap))))
        self._device_state = True

    def move_discard(self):
        nat.check(nat.lib().lg_move_discard(self._sync()))

    def mc_run(self, seed, attempt0, nattempts, beta, delta):
        """`nattempts` Translate attempts of MonteCarlo::propagate on the
        device (monte_carlo.rs:289-330); returns the accepted count."""
        acc = ctypes.c_uint64()
        nat.check(nat.lib().lg_mc_run(self._sync(), int(seed), int(attempt0), int(nattempts), float(beta),
                                      float(delta), ctypes.byref(acc)))
        self._device_state = True
        return acc.value


def mc_run_batch(systems, seeds, attempt0, nattempts, beta, delta):
    """lg_mc_run for independent chains in one launch; per-chain arguments are
    sequences (or scalars, broadcast).  Returns the accepted counts."""
    count = len(systems)

    def arr(x, dtype):
        a = np.broadcast_to(np.asarray(x, dtype=dtype), (count,))
        return np.ascontiguousarray(a)

    handles = (ctypes.c_void_p * count)(*[s._sync().value for s in systems])
    seeds, a0, na = arr(seeds, np.uint64), arr(attempt0, np.uint64), arr(nattempts, np.uint64)
    beta, delta = arr(beta, float), arr(delta, float)
    acc = np.zeros(count, dtype=np.uint64)
    u64 = lambda a: a.ctypes.data_as(nat.U64P)
    nat.check(nat.lib().lg_mc_run_batch(handles, count, u64(seeds), u64(a0), u64(na), nat.dptr(beta),
                                        nat.dptr(delta), u64(acc)))
    for s in systems:
        s._device_state = True
    return acc


# ---------------------------------------------------------------------------------
# molecular dynamics (lumol-sim)
# ---------------------------------------------------------------------------------
class VelocityVerlet:
    """integrators.rs:22-70"""
    code = nat.LG_INTEGRATOR_VELOCITY_VERLET

    def __init__(self, timestep):
        self.timestep = float(timestep)


class LeapFrog:
    """integrators.rs:128-169"""
    code = nat.LG_INTEGRATOR_LEAP_FROG

    def __init__(self, timestep):
        self.timestep = float(timestep)


class BerendsenThermostat:
    """thermostats.rs:91-121"""

    def __init__(self, temperature, tau):
        if temperature < 0:
            raise ValueError("The temperature must be positive in thermostats.")
        if tau < 1.0:
            raise ValueError("The timestep must be larger than 1 in berendsen thermostat.")
        self.temperature = float(temperature)
        self.tau = float(tau)


class MolecularDynamics:
    """molecular_dynamics.rs:11-83: integrator + optional thermostat +
    controls, applied in that order each step.  VelocityVerlet and LeapFrog
    with no thermostat or a Berendsen one run fused on the device; the other
    integrators, thermostats and controls (lumol_amd.propagators) step on the
    host around device force evaluations."""

    def __init__(self, integrator):
        self.integrator = integrator
        self.thermostat = None
        self.controls = []

    @classmethod
    def new(cls, timestep):
        """MolecularDynamics::new: a VelocityVerlet integrator."""
        return cls(VelocityVerlet(timestep))

    def set_thermostat(self, thermostat):
        self.thermostat = thermostat

    def add_control(self, control):
        self.controls.append(control)


# ---------------------------------------------------------------------------------
# Monte Carlo (lumol-sim/src/mc)
# ---------------------------------------------------------------------------------
class Translate:
    """translate.rs:22-148: displaces one random molecule by U(+-delta/sqrt(3))
    per axis."""

    def __init__(self, delta):
        if delta <= 0:
            raise ValueError("delta must be positive in Translate move")
        self.delta = float(delta)


class MonteCarlo:
    """monte_carlo.rs:17-345: Metropolis sampling at `temperature` (K) over
    moves picked by frequency.  A single Translate move without amplitude
    updates runs as the device chain (random numbers from Philox4x32-10 keyed
    by `seed`, one counter per attempt); other move sets (Rotate, Resize,
    several moves, amplitude updates) are driven from the host with device
    energy changes (lumol_amd.propagators)."""

    def __init__(self, temperature, seed=0):
        if not temperature > 0:
            raise ValueError(f"Monte Carlo temperature must be positive, got {temperature}")
        self.temperature = float(temperature)
        self.beta = 1.0 / (K_BOLTZMANN * self.temperature)
        self.seed = int(seed)
        self.moves = []  # (move, frequency, target_acceptance)
        self.update_frequency = 0
        self.attempts = 0
        self.accepted = 0

    def add(self, move, frequency=1.0, target_acceptance=None):
        if target_acceptance is not None and not 0.0 < target_acceptance < 1.0:
            raise ValueError(f"target acceptance ratio must be between 0 and 1, got {target_acceptance}")
        self.moves.append((move, float(frequency), target_acceptance))

    def set_amplitude_update_frequency(self, frequency):
        self.update_frequency = int(frequency)

    @property
    def move(self):
        return self.moves[0][0] if self.moves else None

    def device_chain(self):
        """True when the run is the device chain kernel: one Translate move."""
        return (len(self.moves) == 1 and isinstance(self.moves[0][0], Translate)
                and (self.update_frequency == 0 or self.moves[0][2] is None))


class Simulation:
    """simulations.rs:54-159: runs `nsteps` propagator steps (MD steps, or
    Monte Carlo attempts)."""

    def __init__(self, propagator):
        self.propagator = propagator
        self._setup_for = None
        self._host = None

    def _md_mode(self):
        from . import propagators as P
        md = self.propagator
        fused_int = isinstance(md.integrator, (VelocityVerlet, LeapFrog))
        th = md.thermostat
        if fused_int and (th is None or isinstance(th, BerendsenThermostat)) and not md.controls:
            return "device"
        if fused_int:
            return "hybrid"
        if not isinstance(md.integrator, (P.Verlet, P.BerendsenBarostat, P.AnisoBerendsenBarostat)):
            raise ValueError(f"unknown integrator {type(md.integrator).__name__}")
        return "host"

    def setup(self, system):
        h = system._sync()
        md = self.propagator
        if isinstance(md, MonteCarlo):
            # monte_carlo.rs:247-280: Translate moves whole molecules
            system.simulated_degrees_of_freedom = "molecules"
            system.external_temperature = md.temperature  # simulations.rs:70-75
            system.cache_init()
            self._setup_for = (id(system), h.value)
            self._host = None
            return
        system.simulated_degrees_of_freedom = "particles"  # molecular_dynamics.rs:53-57
        system.external_temperature = None
        mode = self._md_mode()
        if mode == "host":
            md.integrator.setup(system)
        else:
            # the device integrator; a Berendsen thermostat runs on the device
            # right after it, other thermostats and the controls on the host
            th = md.thermostat
            T, tau = (th.temperature, th.tau) if isinstance(th, BerendsenThermostat) else (-1.0, 1.0)
            nat.check(nat.lib().lg_md_setup(h, md.integrator.code, md.integrator.timestep, T, tau))
        self._setup_for = (id(system), h.value)

    def run(self, system, nsteps):
        h = system._sync()
        if self._setup_for != (id(system), h.value):
            self.setup(system)
        p = self.propagator
        if isinstance(p, MonteCarlo):
            if not p.moves:
                raise ValueError("no move added to the Monte Carlo propagator")
            if p.device_chain():
                p.accepted += system.mc_run(p.seed, p.attempts, int(nsteps), p.beta, p.moves[0][0].delta)
                p.attempts += int(nsteps)
            else:
                self._run_mc_host(system, int(nsteps))
            return
        mode = self._md_mode()
        if mode == "device":
            nat.check(nat.lib().lg_md_run(h, int(nsteps)))
            system._device_state = True
            return
        th = p.thermostat if not isinstance(p.thermostat, BerendsenThermostat) or mode == "host" else None
        if mode == "host" and isinstance(p.thermostat, BerendsenThermostat):
            from .propagators import HostBerendsen
            th = HostBerendsen(p.thermostat.temperature, p.thermostat.tau)
        for _ in range(int(nsteps)):
            if mode == "hybrid":
                nat.check(nat.lib().lg_md_run(system._sync(), 1))
                system._device_state = True
            else:
                p.integrator.integrate(system)
            if th is not None:
                th.apply(system)
            for c in p.controls:
                c.control(system)

    def _run_mc_host(self, system, nsteps):
        """MonteCarlo::propagate (monte_carlo.rs:289-330) on the host."""
        from . import propagators as P
        mc = self.propagator
        if self._host is None:
            counters = []
            for move, freq, target in mc.moves:
                hm = P.HostTranslate(move.delta) if isinstance(move, Translate) else move
                counters.append(P.MoveCounter(hm, target))
            freqs = np.array([f for _, f, _ in mc.moves], dtype=float)
            freqs = freqs / freqs.sum()
            cum = np.cumsum(freqs)
            cum[-1] = 1.0
            self._host = (counters, cum, np.random.default_rng(mc.seed))
        counters, cum, rng = self._host
        mols = system.molecules()
        for _ in range(nsteps):
            u = rng.random()
            k = int(np.searchsorted(cum, u, side="left"))
            c = counters[min(k, len(counters) - 1)]
            mv = c.move
            if not mv.prepare(system, rng, mols):
                continue
            cost = mv.cost(system, mc.beta)
            if cost <= 0.0 or rng.random() < math.exp(-cost):
                mv.apply(system)
                c.accept()
                mc.accepted += 1
            else:
                mv.restore(system)
                c.reject()
            mc.attempts += 1
            if c.attempted == mc.update_frequency:
                mv.update_amplitude(c.scaling_factor(), system)
                c.attempted = c.accepted = 0
        system._device_state = True
