"""The reference's input files (lumol-input): TOML simulation files with their
interaction files, and the XYZ and PDB configurations they point to.

  Input(path).read() -> Config(system, simulation, nsteps)   lumol-input/src/simulations/mod.rs:69-110
  InteractionsInput(path or table).read(system)              lumol-input/src/interactions/mod.rs:22-63
  [[systems]] file, cell, guess_bonds, potentials,
              velocities = {init = "T"}                      lumol-input/src/simulations/system.rs:15-180
  [pairs] / [bonds] / [angles] / [dihedrals] / [coulomb]
  / [charges], restrictions, table computations              lumol-input/src/interactions/*.rs
  [[simulations]] nsteps, propagator                          lumol-input/src/simulations/{md,mc,propagator}.rs

Structure files are read here (the reference goes through chemfiles 0.10, a
C++ library that is not part of the reference checkout): XYZ (name x y z
lines after a count and a comment line) and PDB (CRYST1, ATOM/HETATM and
CONECT records).  Masses come from lumol's table of atomic masses
(lumol-core/src/sys/config/mass.rs).  Bonds guessed from distances restate
chemfiles' published criterion (the VMD one): two atoms are bonded when
0.03 A < r < 0.6 (r_vdw,i + r_vdw,j), with Bondi van der Waals radii.

Velocities drawn by `velocities = {init = "T"}` follow BoltzmannVelocities
(lumol-sim/src/velocities.rs:33-84): N(0, kT/m) per component, translation
and rotation removed, rescaled to T exactly.  The reference draws from a
xorshift stream that no test pins; this reader uses numpy's generator with a
fixed seed: same distribution, other numbers.
"""
import math
import os

import numpy as np

try:
    import tomllib as _toml  # Python 3.11+
except ImportError:  # pragma: no cover - the image's Python is 3.10
    import tomli as _toml

from . import unitconv
from .consts import FOUR_PI_EPSILON_0, K_BOLTZMANN
from .system import (ATOMIC_MASSES, BerendsenThermostat, BornMayerHuggins, Buckingham, CosineHarmonic, Ewald,
                     Gaussian, Harmonic, LeapFrog, LennardJones, Mie, MolecularDynamics, MonteCarlo, Morse,
                     NullPotential, PairInteraction, PairRestriction, Simulation, System, TableComputation,
                     Torsion, Translate, UnitCell, VelocityVerlet, Wolf)


FRAC_2_SQRT_PI = 2.0 / math.sqrt(math.pi)


class InputError(ValueError):
    """An input file the reference would refuse, or that names something the
    device backend does not run."""


# --- helpers of lumol-input/src/extract.rs ------------------------------------------
def _get(table, key, context, kind=None):
    if key not in table:
        raise InputError(f"missing '{key}' key in {context}")
    v = table[key]
    if kind == "str" and not isinstance(v, str):
        raise InputError(f"'{key}' must be a string in {context}")
    if kind == "number" and (isinstance(v, bool) or not isinstance(v, (int, float))):
        raise InputError(f"'{key}' must be a number in {context}")
    if kind == "uint" and (isinstance(v, bool) or not isinstance(v, int) or v < 0):
        raise InputError(f"'{key}' must be a positive integer in {context}")
    if kind == "table" and not isinstance(v, dict):
        raise InputError(f"'{key}' must be a table in {context}")
    if kind == "slice" and not isinstance(v, list):
        raise InputError(f"'{key}' must be an array in {context}")
    return v


def _unit(table, key, context):
    try:
        return unitconv.from_str(_get(table, key, context, "str"))
    except unitconv.UnitError as e:
        raise InputError(f"in {context}, '{key}': {e}") from None


def _type(table, context):
    return _get(table, "type", context, "str")


def _validate(config):
    # lumol-input/src/lib.rs:78-95
    if "input" not in config:
        raise InputError("missing 'input' table")
    if "version" not in config["input"]:
        raise InputError("missing 'version' key in 'input' table")
    version = config["input"]["version"]
    if isinstance(version, bool) or not isinstance(version, int):
        raise InputError("'input.version' must be an integer")
    if version != 1:
        raise InputError(f"can only read version 1 of input, got version {version}")


def _load_toml(path):
    try:
        with open(path, "rb") as f:
            return _toml.load(f)
    except OSError as e:
        raise InputError(f"can not read {path}: {e}") from None
    except _toml.TOMLDecodeError as e:
        raise InputError(f"TOML error in {path}: {e}") from None


# --- structure files --------------------------------------------------------------------
# Bondi van der Waals radii (A) of the elements bond guessing meets
VDW_RADII = {"H": 1.2, "He": 1.4, "C": 1.7, "N": 1.55, "O": 1.52, "F": 1.47, "Ne": 1.54, "Na": 2.27,
             "Mg": 1.73, "Si": 2.1, "P": 1.8, "S": 1.8, "Cl": 1.75, "Ar": 1.88, "K": 2.75, "Kr": 2.02,
             "Xe": 2.16}


def read_xyz(path):
    """names, positions (N x 3) of an XYZ file."""
    with open(path) as f:
        lines = f.read().splitlines()
    try:
        n = int(lines[0].split()[0])
    except (IndexError, ValueError):
        raise InputError(f"{path}: the first line must hold the number of atoms") from None
    if len(lines) < 2 + n:
        raise InputError(f"{path}: expected {n} atoms")
    names, pos = [], []
    for line in lines[2:2 + n]:
        t = line.split()
        if len(t) < 4:
            raise InputError(f"{path}: malformed atom line '{line}'")
        names.append(t[0])
        pos.append([float(t[1]), float(t[2]), float(t[3])])
    return names, np.array(pos, dtype=float).reshape(-1, 3)


def read_pdb(path):
    """names, positions, cell lengths (or None), bonds of a PDB file.  The
    name is the element column (77-78), or the atom name (13-16) without
    digits when that is empty; bonds come from CONECT records."""
    names, pos, serials, bonds = [], [], {}, set()
    cell = None
    with open(path) as f:
        for line in f:
            rec = line[:6].strip()
            if rec == "CRYST1":
                a, b, c = float(line[6:15]), float(line[15:24]), float(line[24:33])
                angles = [float(line[33:40]), float(line[40:47]), float(line[47:54])]
                if any(abs(x - 90.0) > 1e-6 for x in angles):
                    raise InputError(f"{path}: triclinic cells are not supported by the device backend")
                cell = (a, b, c)
            elif rec in ("ATOM", "HETATM"):
                serial = int(line[6:11])
                element = line[76:78].strip() if len(line) > 76 else ""
                name = element or "".join(ch for ch in line[12:16].strip() if not ch.isdigit())
                serials[serial] = len(names)
                names.append(name)
                pos.append([float(line[30:38]), float(line[38:46]), float(line[46:54])])
            elif rec == "CONECT":
                fields = [line[i:i + 5] for i in range(6, 31, 5)]
                ids = [int(x) for x in fields if x.strip()]
                for other in ids[1:]:
                    i, j = serials[ids[0]], serials[other]
                    if i != j:
                        bonds.add((min(i, j), max(i, j)))
    if cell is not None and cell == (0.0, 0.0, 0.0):
        cell = None
    return names, np.array(pos, dtype=float).reshape(-1, 3), cell, sorted(bonds)


def guess_bonds(names, positions, cell=None):
    """Bonds from distances: 0.03 A < r < 0.6 (r_vdw,i + r_vdw,j) with the
    minimum image in a periodic cell."""
    pos = np.asarray(positions, dtype=float)
    radii = np.array([VDW_RADII.get(x, 0.0) for x in names])
    n = len(names)
    bonds = []
    L = np.asarray(cell, dtype=float) if cell is not None else None
    for i in range(n - 1):
        d = pos[i + 1:] - pos[i]
        if L is not None:
            d -= np.round(d / L) * L
        r = np.sqrt((d * d).sum(1))
        lim = 0.6 * (radii[i] + radii[i + 1:])
        for k in np.nonzero((r > 0.03) & (r < lim))[0]:
            bonds.append((i, i + 1 + int(k)))
    return bonds


def _contiguous_order(n, bonds):
    """Atom order that makes every bonded component a contiguous range (the
    reference's System::add_bond moves atoms so, configuration.rs:150-200):
    components in the order of their first atom, atoms in index order."""
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for i, j in bonds:
        a, b = find(i), find(j)
        if a != b:
            parent[max(a, b)] = min(a, b)
    groups = {}
    for i in range(n):
        groups.setdefault(find(i), []).append(i)
    return [i for root in sorted(groups) for i in groups[root]]


def read_structure(path, cell=None, guess=False):
    """A System (no potentials yet) from an XYZ or PDB file; `cell` (lengths)
    overrides the file's."""
    ext = os.path.splitext(path)[1].lower()
    file_cell, bonds = None, []
    if ext == ".xyz":
        names, pos = read_xyz(path)
    elif ext == ".pdb":
        names, pos, file_cell, bonds = read_pdb(path)
    else:
        raise InputError(f"unknown structure file format '{ext}' ({path})")
    lengths = cell if cell is not None else file_cell
    if guess:
        bonds = guess_bonds(names, pos, lengths if lengths and all(x > 0 for x in lengths) else None)
    order = _contiguous_order(len(names), bonds)
    if order != list(range(len(names))):
        inv = {old: new for new, old in enumerate(order)}
        names = [names[i] for i in order]
        pos = pos[order]
        bonds = sorted((min(inv[i], inv[j]), max(inv[i], inv[j])) for i, j in bonds)
    uc = UnitCell.infinite() if lengths is None or all(x == 0 for x in lengths) else UnitCell.ortho(*lengths)
    masses = [ATOMIC_MASSES.get(x, 0.0) for x in names]
    s = System.from_arrays(names, pos, uc, masses=masses)
    for i, j in bonds:
        s.add_bond(i, j)
    return s


# --- interactions (lumol-input/src/interactions) ------------------------------------------
def _restriction(table):
    # lumol-input/src/interactions/mod.rs:65-102
    if "restriction" not in table:
        return None
    r = table["restriction"]
    if isinstance(r, str):
        names = {"none": PairRestriction.NONE,
                 "intramolecular": PairRestriction.INTRA_MOLECULAR, "IntraMolecular": PairRestriction.INTRA_MOLECULAR,
                 "intra-molecular": PairRestriction.INTRA_MOLECULAR,
                 "intermolecular": PairRestriction.INTER_MOLECULAR, "InterMolecular": PairRestriction.INTER_MOLECULAR,
                 "inter-molecular": PairRestriction.INTER_MOLECULAR,
                 "exclude12": PairRestriction.EXCLUDE12, "exclude13": PairRestriction.EXCLUDE13,
                 "exclude14": PairRestriction.EXCLUDE14}
        if r == "scale14":
            raise InputError("'scale14' restriction must be a table")
        if r not in names:
            raise InputError(f"Unknown restriction '{r}'")
        return names[r]
    if isinstance(r, dict):
        if list(r) != ["scale14"]:
            raise InputError("Restriction table must be 'scale14'")
        if not isinstance(r["scale14"], float):
            raise InputError("'scale14' parameter must be a float")
        return PairRestriction.scale14(r["scale14"])
    raise InputError("Restriction must be a table or a string")


def _pair_potential(t):
    # lumol-input/src/interactions/pairs.rs:172-184, potentials.rs
    kind = _type(t, "pair potential")
    if kind == "null":
        return NullPotential()
    if kind == "harmonic":
        return Harmonic(k=_unit(t, "k", "harmonic potential"), x0=_unit(t, "x0", "harmonic potential"))
    if kind == "lj":
        return LennardJones(sigma=_unit(t, "sigma", "Lennard-Jones potential"),
                            epsilon=_unit(t, "epsilon", "Lennard-Jones potential"))
    if kind == "buckingham":
        return Buckingham(a=_unit(t, "A", "Buckingham potential"), c=_unit(t, "C", "Buckingham potential"),
                          rho=_unit(t, "rho", "Buckingham potential"))
    if kind == "born":
        c = "Born-Mayer-Huggins potential"
        return BornMayerHuggins(a=_unit(t, "A", c), c=_unit(t, "C", c), d=_unit(t, "D", c),
                                sigma=_unit(t, "sigma", c), rho=_unit(t, "rho", c))
    if kind == "morse":
        return Morse(a=_unit(t, "A", "Morse potential"), x0=_unit(t, "x0", "Morse potential"),
                     depth=_unit(t, "depth", "Morse potential"))
    if kind == "gaussian":
        b = _unit(t, "B", "Gaussian potential")
        if b <= 0:
            raise InputError("'B' parameter has to be positive in Gaussian potential")
        return Gaussian(a=_unit(t, "A", "Gaussian potential"), b=b)
    if kind == "mie":
        m = _get(t, "m", "Mie potential", "number")
        n = _get(t, "n", "Mie potential", "number")
        if not m < n:
            raise InputError("The repulsive exponent n has to be larger than the attractive exponent m")
        return Mie(_unit(t, "sigma", "Mie potential"), _unit(t, "epsilon", "Mie potential"), float(n), float(m))
    raise InputError(f"unknown potential type '{kind}'")


def _bonded_potential(t, what):
    # pairs.rs:186-193 (bonds), angles.rs:60-89 (angles, dihedrals)
    allowed = {"bond": ("null", "harmonic", "morse"),
               "angle": ("null", "harmonic", "cosine-harmonic", "morse"),
               "dihedral": ("null", "harmonic", "cosine-harmonic", "torsion", "morse")}[what]
    kind = _type(t, f"{what} potential")
    if kind not in allowed:
        raise InputError(f"unknown potential type '{kind}'")
    if kind == "null":
        return NullPotential()
    if kind == "harmonic":
        return Harmonic(k=_unit(t, "k", "harmonic potential"), x0=_unit(t, "x0", "harmonic potential"))
    if kind == "morse":
        return Morse(a=_unit(t, "A", "Morse potential"), x0=_unit(t, "x0", "Morse potential"),
                     depth=_unit(t, "depth", "Morse potential"))
    if kind == "cosine-harmonic":
        return CosineHarmonic(k=_unit(t, "k", "cosine harmonic potential"),
                              x0=_unit(t, "x0", "cosine harmonic potential"))
    n = _get(t, "n", "torsion potential", "uint")
    return Torsion(k=_unit(t, "k", "torsion potential"), delta=_unit(t, "delta", "torsion potential"), n=n)


def ewald_with_accuracy(cutoff, accuracy, system):
    """Ewald::with_accuracy (ewald.rs:312-351): alpha and kmax for a target
    relative accuracy of the energy, from the charges and the cell."""
    if cutoff < 0.0:
        raise InputError("the cutoff can not be negative in Ewald")
    if accuracy < 0.0:
        raise InputError("accuracy can not be negative in Ewald")
    q2 = 0.0
    for q in system.charges:
        q2 += q * q
    q2 /= FOUR_PI_EPSILON_0
    natoms = float(system.size())
    a, b, c = system.cell.lengths
    alpha = accuracy * math.sqrt(natoms * cutoff * a * b * c) / (2.0 * q2)
    alpha = (1.35 - 0.15 * math.log(accuracy)) / cutoff if alpha >= 1.0 else math.sqrt(-math.log(alpha)) / cutoff
    min_length = min(a, b, c)

    def error(kmax):
        arg = math.pi * kmax / (alpha * min_length)
        return FRAC_2_SQRT_PI * q2 * alpha / min_length / math.sqrt(kmax * natoms) * math.exp(-arg * arg)

    kmax = 1
    while error(float(kmax)) > accuracy:
        kmax += 1
    return Ewald(cutoff, kmax, alpha)


class InteractionsInput:
    """lumol-input/src/interactions/mod.rs:22-63: a TOML file (or table) of
    interactions read into a System."""

    def __init__(self, source):
        if isinstance(source, dict):
            self.config = source
        else:
            self.config = _load_toml(source)
            _validate(self.config)

    def read(self, system):
        self._pairs(system)
        self._bonded(system, "bonds", 2, system.set_bond_potential, "bond")
        self._bonded(system, "angles", 3, system.set_angle_potential, "angle")
        self._bonded(system, "dihedrals", 4, system.set_dihedral_potential, "dihedral")
        self._charges(system)  # before coulomb (mod.rs:56-58)
        self._coulomb(system)
        return system

    def _pairs(self, system):
        # pairs.rs:55-137
        pairs = self.config.get("pairs")
        if pairs is None:
            return
        if not isinstance(pairs, dict):
            raise InputError("the 'pairs' section must be a table")
        glob = self.config.get("global", {})
        if not isinstance(glob, dict):
            raise InputError("'global' section must be a table")
        g_cutoff = glob.get("cutoff")
        g_tail = glob.get("tail_correction")
        if g_tail is not None and not isinstance(g_tail, bool):
            raise InputError("the 'tail_correction' section must be a boolean value")
        for key, t in pairs.items():
            atoms = key.split("-")
            if len(atoms) != 2:
                raise InputError(f"expected two atoms for pair potential, got {len(atoms)} ({atoms})")
            if not isinstance(t, dict):
                raise InputError(f"pair potential associated with {key} must be a table")
            pot = _pair_potential(t)
            if "computation" in t:
                comp = t["computation"]
                if not isinstance(comp, dict):
                    raise InputError("'computation' section must be a table")
                if len(comp) != 1:
                    raise InputError("Missing computation type in computation table")
                (name, table), = comp.items()
                if name != "table":
                    raise InputError(f"Unknown computation type '{name}'")
                if not isinstance(table, dict):
                    raise InputError("'table' key in computation must be a TOML table")
                pot = TableComputation(pot, _get(table, "n", "table computation", "uint"),
                                       _unit(table, "max", "table computation"))
            cutoff = t.get("cutoff", g_cutoff)
            if cutoff is None:
                raise InputError("missing 'cutoff' value for pair potential")
            if isinstance(cutoff, str):
                inter = PairInteraction(pot, unitconv.from_str(cutoff))
            elif isinstance(cutoff, dict):
                if "shifted" not in cutoff:
                    raise InputError("'cutoff' table can only contain 'shifted' key")
                if not isinstance(cutoff["shifted"], str):
                    raise InputError("'cutoff.shifted' value must be a string")
                inter = PairInteraction.shifted(pot, unitconv.from_str(cutoff["shifted"]))
            else:
                raise InputError("'cutoff' must be a string or a table")
            tail = t.get("tail_correction", g_tail)
            if tail is not None and not isinstance(tail, bool):
                raise InputError("the 'tail_correction' section must be a boolean value")
            if tail:
                inter.enable_tail_corrections()
            r = _restriction(t)
            if r is not None:
                inter.set_restriction(r)
            system.set_pair_potential((atoms[0], atoms[1]), inter)

    def _bonded(self, system, section, natoms, setter, what):
        table = self.config.get(section)
        if table is None:
            return
        if not isinstance(table, dict):
            raise InputError(f"the '{section}' section must be a table")
        for key, t in table.items():
            atoms = key.split("-")
            if len(atoms) != natoms:
                raise InputError(f"expected {natoms} atoms for {what} potential, got {len(atoms)} ({atoms})")
            if not isinstance(t, dict):
                raise InputError(f"{what} potential associated with {key} must be a table")
            setter(tuple(atoms), _bonded_potential(t, what))

    def _charges(self, system):
        # coulomb.rs:50-90
        charges = self.config.get("charges")
        if charges is None:
            return
        if not isinstance(charges, dict):
            raise InputError("the 'charges' section must be a table")
        q = np.array(system.charges, dtype=float)
        for name, value in charges.items():
            if isinstance(value, bool) or not isinstance(value, (int, float)):
                raise InputError("charges must be numbers")
            for i, x in enumerate(system.names):
                if x == name:
                    q[i] = float(value)
        system.charges = q
        system.particles_changed()

    def _coulomb(self, system):
        # coulomb.rs:14-48
        coul = self.config.get("coulomb")
        if coul is None:
            return
        if not isinstance(coul, dict):
            raise InputError("the 'coulomb' section must be a table")
        solvers = [k for k in coul if k != "restriction"]
        if len(solvers) != 1:
            raise InputError(f"got more than one coulombic solver: {' and '.join(solvers)}")
        key = solvers[0]
        t = coul[key]
        if not isinstance(t, dict):
            raise InputError(f"coulombic solver '{key}' must be a table")
        if key == "wolf":
            pot = Wolf(_unit(t, "cutoff", "Wolf coulombic potential"))
        elif key == "ewald":
            # potentials.rs:164-188
            cutoff = _unit(t, "cutoff", "Ewald coulombic potential")
            if "accuracy" in t:
                if "kmax" in t or "alpha" in t:
                    raise InputError("can not have both accuracy and kmax/alpha in Ewald coulombic potential")
                pot = ewald_with_accuracy(cutoff, float(_get(t, "accuracy", "Ewald coulombic potential", "number")),
                                          system)
            else:
                kmax = _get(t, "kmax", "Ewald coulombic potential", "uint")
                alpha = _unit(t, "alpha", "Ewald coulombic potential") if "alpha" in t else None
                pot = Ewald(cutoff, kmax, alpha)
        else:
            raise InputError(f"unknown coulomb solver '{key}'")
        r = _restriction(coul)
        if r is not None:
            pot.set_restriction(r)
        system.set_coulomb_potential(pot)


# --- velocities (lumol-sim/src/velocities.rs, md/controls.rs) ----------------------------
def _cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


def remove_translation(masses, velocities):
    """RemoveTranslation (controls.rs:27-43)."""
    total = 0.0
    for m in masses:
        total += m
    com_v = np.zeros(3)
    for m, v in zip(masses, velocities):
        com_v += v * m / total
    return velocities - com_v


def _inverse3(m):
    # Matrix3::inverse (types/matrix.rs:212-229)
    det = (m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2])
           - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
           + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))
    if not abs(det) > 1e-30:
        raise InputError("The matrix is not invertible!")
    inv = 1.0 / det
    r = np.zeros((3, 3))
    r[0][0] = (m[1][1] * m[2][2] - m[2][1] * m[1][2]) * inv
    r[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv
    r[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv
    r[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * inv
    r[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv
    r[1][2] = (m[1][0] * m[0][2] - m[0][0] * m[1][2]) * inv
    r[2][0] = (m[1][0] * m[2][1] - m[2][0] * m[1][1]) * inv
    r[2][1] = (m[2][0] * m[0][1] - m[0][0] * m[2][1]) * inv
    r[2][2] = (m[0][0] * m[1][1] - m[1][0] * m[0][1]) * inv
    return r


def remove_rotation(masses, positions, velocities):
    """RemoveRotation (controls.rs:46-76): subtract the rigid rotation
    omega = I^-1 L about the centre of mass, with the reference's inertia
    matrix (-sum m d (x) d, plus its trace on the diagonal)."""
    total = 0.0
    com = np.zeros(3)
    for m, p in zip(masses, positions):
        total += m
        com += m * p
    com = com / total
    moment = np.zeros(3)
    inertia = np.zeros((3, 3))
    for m, p, v in zip(masses, positions, velocities):
        d = p - com
        moment += m * _cross(d, v)
        inertia += -m * np.outer(d, d)
    tr = inertia[0][0] + inertia[1][1] + inertia[2][2]
    for k in range(3):
        inertia[k][k] += tr
    inv = _inverse3(inertia)
    angular = inv @ moment
    out = np.array(velocities, dtype=float)
    for i, p in enumerate(positions):
        out[i] -= _cross(p - com, angular)
    return out


def boltzmann_velocities(system, temperature, seed=42):
    """BoltzmannVelocities::init (velocities.rs:56-67) with numpy's
    generator: draws, removes translation and rotation, rescales to T."""
    rng = np.random.default_rng(seed)
    m = np.asarray(system.masses, dtype=float)
    if (m <= 0).any():
        raise InputError("every particle needs a positive mass to draw velocities")
    sigma = math.sqrt(K_BOLTZMANN * temperature)
    v = np.sqrt(1.0 / m)[:, None] * rng.normal(0.0, sigma, size=(len(m), 3))
    v = remove_translation(m, v)
    v = remove_rotation(m, system.positions, v)
    ke = 0.0
    for mi, vi in zip(m, v):
        ke += 0.5 * mi * float(vi @ vi)
    t_now = 2.0 * ke / (3 * len(m) * K_BOLTZMANN)
    system.velocities = v * math.sqrt(temperature / t_now)
    return system


# --- simulations ------------------------------------------------------------------------------
class Config:
    """What Input::read returns (lumol-input/src/simulations/mod.rs:31-39)."""

    def __init__(self, system, simulation, nsteps):
        self.system = system
        self.simulation = simulation
        self.nsteps = nsteps


def _md_propagator(t):
    # lumol-input/src/simulations/md.rs:14-196
    from . import propagators as P
    dt = _unit(t, "timestep", "molecular dynamics propagator")
    integrator = VelocityVerlet(dt)
    if "integrator" in t:
        it = t["integrator"]
        if not isinstance(it, dict):
            raise InputError("'integrator' must be a table in molecular dynamics")
        kind = _type(it, "integrator")
        if kind == "VelocityVerlet":
            integrator = VelocityVerlet(dt)
        elif kind == "LeapFrog":
            integrator = LeapFrog(dt)
        elif kind == "Verlet":
            integrator = P.Verlet(dt)
        elif kind == "BerendsenBarostat":
            integrator = P.BerendsenBarostat(dt, _unit(it, "pressure", "Berendsen barostat"),
                                             float(_get(it, "timestep", "Berendsen barostat", "number")))
        elif kind == "AnisoBerendsenBarostat":
            c = "anisotropic Berendsen barostat"
            integrator = P.AnisoBerendsenBarostat.hydrostatic(dt, _unit(it, "pressure", c),
                                                              float(_get(it, "timestep", c, "number")))
        else:
            raise InputError(f"unknown integrator '{kind}'")
    md = MolecularDynamics(integrator)
    if "thermostat" in t:
        th = t["thermostat"]
        if not isinstance(th, dict):
            raise InputError("'thermostat' must be a table in molecular dynamics")
        kind = _type(th, "thermostat")
        if kind == "Berendsen":
            md.set_thermostat(BerendsenThermostat(_unit(th, "temperature", "Berendsen thermostat"),
                                                  float(_get(th, "timestep", "Berendsen thermostat", "number"))))
        elif kind == "Rescale":
            T = _unit(th, "temperature", "rescale thermostat")
            tol = None
            if "tolerance" in th:
                if not isinstance(th["tolerance"], str):
                    raise InputError("'tolerance' must be a string rescale thermostat")
                tol = unitconv.from_str(th["tolerance"])
            md.set_thermostat(P.RescaleThermostat(T, tol))
        elif kind == "CSVR":
            md.set_thermostat(P.CSVRThermostat(_unit(th, "temperature", "CSVR thermostat"),
                                               float(_get(th, "timestep", "CSVR thermostat", "number"))))
        else:
            raise InputError(f"unknown thermostat type '{kind}'")
    if "controls" in t:
        controls = t["controls"]
        if not isinstance(controls, list):
            raise InputError("'controls' must be an array of tables in molecular dynamics")
        kinds = {"RemoveTranslation": P.RemoveTranslation, "RemoveRotation": P.RemoveRotation,
                 "Rewrap": P.Rewrap}
        for c in controls:
            if not isinstance(c, dict):
                raise InputError("'controls' must be an array of tables in molecular dynamics")
            kind = _type(c, "control")
            if kind not in kinds:
                raise InputError(f"unknown control '{kind}'")
            every = _get(c, "every", f"{kind} control", "uint") if "every" in c else 1
            md.add_control(kinds[kind](every))
    return md


def _mc_propagator(t):
    # lumol-input/src/simulations/mc.rs:15-125
    from . import propagators as P
    temperature = _unit(t, "temperature", "Monte Carlo propagator")
    moves = _get(t, "moves", "Monte Carlo propagator", "slice")
    has_update = "update_frequency" in t
    mc = MonteCarlo(temperature)
    for mv in moves:
        if not isinstance(mv, dict):
            raise InputError("All moves must be tables in Monte Carlo")
        freq = float(_get(mv, "frequency", "Monte Carlo move", "number")) if "frequency" in mv else 1.0
        target = None
        if "target_acceptance" in mv:
            target = float(_get(mv, "target_acceptance", "Monte Carlo move", "number"))
            if not has_update:
                raise InputError("No 'update_frequency' found. Please specify 'update_frequency' in "
                                 "combination with 'target_acceptance'")
            if not 0.0 <= target <= 1.0:
                raise InputError("'target_acceptance' has to be between 0.0 and 1.0")
        kind = _type(mv, "Monte Carlo move")
        if "molecule" in mv:
            raise InputError(f"{kind} moves restricted to one kind of molecule are not supported")
        if kind == "Translate":
            move = Translate(_unit(mv, "delta", "Translate move"))
        elif kind == "Rotate":
            move = P.Rotate(_unit(mv, "delta", "Rotate move"))
        elif kind == "Resize":
            move = P.Resize(_unit(mv, "pressure", "Resize move"), _unit(mv, "delta", "Resize move"))
        else:
            raise InputError(f"unknown Monte Carlo move '{kind}'")
        mc.add(move, freq, target)
    if has_update:
        mc.set_amplitude_update_frequency(_get(t, "update_frequency", "Monte Carlo propagator", "uint"))
    return mc


class Input:
    """A simulation input file (lumol-input/src/simulations/mod.rs:46-110)."""

    def __init__(self, path):
        self.path = os.path.abspath(path)
        self.config = _load_toml(self.path)
        _validate(self.config)

    def _rel(self, name):
        return name if os.path.isabs(name) else os.path.join(os.path.dirname(self.path), name)

    def _system_table(self):
        systems = _get(self.config, "systems", "input file", "slice")
        if not systems:
            raise InputError("'systems' array should contain a system")
        if len(systems) > 1:
            raise InputError("only one system is supported in input file")
        if not isinstance(systems[0], dict):
            raise InputError("'systems' should be an array of tables in input file")
        return systems[0]

    def _cell(self, t):
        # system.rs:77-115
        if "cell" not in t:
            return None
        c = t["cell"]
        if isinstance(c, bool):
            raise InputError("'cell' must be a number or an array in system")
        if isinstance(c, (int, float)):
            return (float(c),) * 3
        if isinstance(c, list):
            if not c:
                return (0.0, 0.0, 0.0)
            if not all(isinstance(x, (int, float)) and not isinstance(x, bool) for x in c):
                raise InputError("values must be numbers in 'cell' array")
            if len(c) == 3:
                return tuple(float(x) for x in c)
            if len(c) == 6:
                if any(abs(float(x) - 90.0) > 1e-12 for x in c[3:]):
                    raise InputError("triclinic cells are not supported by the device backend")
                return tuple(float(x) for x in c[:3])
            raise InputError("'cell' array must have a size of 3 or 6")
        raise InputError("'cell' must be a number or an array in system")

    def read_system(self):
        t = self._system_table()
        path = self._rel(_get(t, "file", "system", "str"))
        guess = t.get("guess_bonds", False)
        if not isinstance(guess, bool):
            raise InputError("'guess_bonds' should be a boolean value in system")
        if "topology" in t:
            raise InputError("separate topology files are not supported")
        system = read_structure(path, cell=self._cell(t), guess=guess)
        pots = t.get("potentials")
        if isinstance(pots, str):
            InteractionsInput(self._rel(pots)).read(system)
        elif isinstance(pots, dict):
            InteractionsInput(pots).read(system)
        elif pots is not None:
            raise InputError("'potentials' must be a string or a table in system")
        vel = t.get("velocities")
        if vel is not None:
            if not isinstance(vel, dict):
                raise InputError("'velocities' must be a table in system")
            if "init" in vel:
                boltzmann_velocities(system, _unit(vel, "init", "velocities initializer"))
        return system

    def _simulation_table(self):
        sims = _get(self.config, "simulations", "input file", "slice")
        if len(sims) != 1:
            raise InputError("only one simulation is supported in the input")
        if not isinstance(sims[0], dict):
            raise InputError("simulations should be tables")
        return sims[0]

    def read_nsteps(self):
        t = self._simulation_table()
        if "nsteps" not in t:
            raise InputError("missing 'nsteps' key in simulation")
        if isinstance(t["nsteps"], bool) or not isinstance(t["nsteps"], int):
            raise InputError("'nsteps' key must be an integer")
        return t["nsteps"]

    def read_simulation(self):
        t = self._simulation_table()
        prop = _get(t, "propagator", "simulation", "table")
        kind = _type(prop, "propagator")
        if kind == "MolecularDynamics":
            return Simulation(_md_propagator(prop))
        if kind == "MonteCarlo":
            return Simulation(_mc_propagator(prop))
        if kind == "Minimization":
            raise InputError("minimization is not supported by the device backend")
        raise InputError(f"unknown propagator type '{kind}'")

    def read(self):
        system = self.read_system()
        simulation = self.read_simulation()
        return Config(system, simulation, self.read_nsteps())
