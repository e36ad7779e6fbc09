"""Synthetic systems of the benchmark configurations and an XYZ reader.

Each builder returns a ``lumol_amd.System`` with seeded, reproducible
coordinates (numpy PCG64 streams), so the GPU path and the oracle see the
same bytes:

* ``argon``  -- LJ argon NVE: FCC lattice with uniform jitter, shifted LJ.
* ``nacl``   -- rocksalt NaCl: Lorentz-Berthelot LJ + Ewald, Berendsen NVT.
* ``spce``   -- flexible SPC/E water on a jittered cubic grid, random
  orientations, harmonic bonds/angles, inter-molecular LJ + Ewald.
* ``kob_andersen`` -- the Kob-Andersen 80:20 binary LJ mixture at
  rho* = 1.2 (sigma_AA = 3.405 A, eps_AA = 0.997 kJ/mol), one cutoff
  2.5 sigma_AA for every pair.
* ``lj_mc_chain`` -- one Monte Carlo chain of LJ argon at rho* = 0.8,
  rc = 2.5 sigma, with its temperature (T* = 1.2) and amplitude (0.3 sigma).

Masses and charges follow the configurations; energies are converted from
kJ/mol and kcal/mol to lumol's internal units (lumol-core/src/units.rs).
"""
import math

import numpy as np

from .system import (ATOMIC_MASSES, BerendsenThermostat, Ewald, Harmonic, K_BOLTZMANN, LennardJones,
                     MolecularDynamics, MonteCarlo, NullPotential, PairInteraction, PairRestriction, Simulation,
                     System, Translate, UnitCell, VelocityVerlet, units)


LJ_SIGMA = 3.405
LJ_EPSILON = units.from_(0.997, "kJ/mol")


def read_xyz(path):
    """XYZ with an optional ``cell: a [b c]`` comment line (tests/nist-spce.rs:24-46)."""
    with open(path) as f:
        lines = f.read().splitlines()
    n = int(lines[0].split()[0])
    cell = None
    parts = lines[1].split()
    if parts and parts[0] == "cell:":
        vals = [float(x) for x in parts[1:4]]
        cell = (vals * 3)[:3] if len(vals) == 1 else vals
    names, pos = [], []
    for line in lines[2:2 + n]:
        t = line.split()
        names.append(t[0])
        pos.append([float(t[1]), float(t[2]), float(t[3])])
    return names, np.array(pos), cell


def maxwell_boltzmann(masses, temperature, rng):
    """Velocities from N(0, kT/m) with the centre-of-mass momentum removed and
    rescaled to `temperature` (semantics of lumol-sim/src/velocities.rs:56-67;
    the rotation removal is omitted for periodic systems)."""
    m = np.asarray(masses, dtype=float)
    v = rng.normal(0.0, math.sqrt(K_BOLTZMANN * temperature), size=(len(m), 3)) / np.sqrt(m)[:, None]
    v -= (m[:, None] * v).sum(0) / m.sum()
    ke = 0.5 * (m[:, None] * v * v).sum()
    t_now = 2.0 * ke / (3 * len(m) * K_BOLTZMANN)
    return v * math.sqrt(temperature / t_now)


def argon(ncell=5, a=5.74, jitter=0.1, seed=1, temperature=120.0, cutoff=8.5):
    """500-atom LJ argon: FCC ncell^3, eps 0.997 kJ/mol, sigma 3.405 A, shifted."""
    rng = np.random.default_rng(seed)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    grid = np.array([[i, j, k] for i in range(ncell) for j in range(ncell) for k in range(ncell)], float)
    pos = ((grid[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a
    pos += rng.uniform(-jitter, jitter, size=pos.shape)
    L = ncell * a
    n = len(pos)
    masses = np.full(n, 39.948)
    vel = maxwell_boltzmann(masses, temperature, rng)
    s = System.from_arrays(["Ar"] * n, pos, UnitCell.cubic(L), masses=masses, velocities=vel)
    lj = LennardJones(sigma=3.405, epsilon=units.from_(0.997, "kJ/mol"))
    s.set_pair_potential(("Ar", "Ar"), PairInteraction.shifted(lj, cutoff))
    return s


def nacl(ncell=8, a=5.64, jitter=0.05, seed=2, temperature=1200.0, cutoff=12.0, kmax=9):
    """Rocksalt NaCl (8 ions per cubic cell), LB-mixed LJ + Ewald(cutoff, kmax)."""
    rng = np.random.default_rng(seed)
    na = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cl = na + np.array([0.5, 0.0, 0.0])
    grid = np.array([[i, j, k] for i in range(ncell) for j in range(ncell) for k in range(ncell)], float)
    names, pos = [], []
    for g in grid:
        for b in na:
            names.append("Na")
            pos.append((g + b) * a)
        for b in cl:
            names.append("Cl")
            pos.append((g + b) * a)
    pos = np.array(pos)
    pos += rng.uniform(-jitter, jitter, size=pos.shape)
    L = ncell * a
    charges = np.array([1.0 if x == "Na" else -1.0 for x in names])
    masses = np.array([ATOMIC_MASSES[x] for x in names])
    vel = maxwell_boltzmann(masses, temperature, rng)
    s = System.from_arrays(names, pos, UnitCell.cubic(L), charges=charges, masses=masses, velocities=vel)
    sig = {"Na": 2.35, "Cl": 4.40}
    eps = {"Na": 0.544, "Cl": 0.419}
    for x, y in (("Na", "Na"), ("Cl", "Cl"), ("Na", "Cl")):
        lj = LennardJones(sigma=0.5 * (sig[x] + sig[y]),
                          epsilon=units.from_(math.sqrt(eps[x] * eps[y]), "kJ/mol"))
        s.set_pair_potential((x, y), PairInteraction(lj, cutoff))
    s.set_coulomb_potential(Ewald(cutoff, kmax))
    return s


def spce(nmol=3000, L=44.8, seed=3, temperature=300.0, cutoff=10.0, kmax=10, jitter=0.2):
    """Flexible SPC/E water: O-H 1 A, H-O-H 109.47 deg, qO -0.8476."""
    rng = np.random.default_rng(seed)
    m = int(math.ceil(nmol ** (1.0 / 3.0)))
    spacing = L / m
    sites = np.array([[i, j, k] for i in range(m) for j in range(m) for k in range(m)], float)
    sites = sites[rng.permutation(len(sites))[:nmol]]
    sites = sites[np.lexsort((sites[:, 2], sites[:, 1], sites[:, 0]))]
    centres = (sites + 0.5) * spacing + rng.uniform(-jitter, jitter, size=sites.shape)
    theta = math.radians(109.47)
    local = np.array([[0.0, 0.0, 0.0],
                      [1.0, 0.0, 0.0],
                      [math.cos(theta), math.sin(theta), 0.0]])
    # one random orientation per molecule (stacked QR: the same numbers as
    # one QR per molecule in order)
    q = np.linalg.qr(rng.normal(size=(nmol, 3, 3)))[0]
    pos = (np.matmul(local[None], np.transpose(q, (0, 2, 1))) + centres[:, None, :]).reshape(-1, 3)
    names = ["O", "H", "H"] * nmol
    base = 3 * np.arange(nmol)
    bonds = [(int(b), int(b) + 1) for b in base] + [(int(b), int(b) + 2) for b in base]
    bonds.sort()
    n = len(names)
    charges = np.array([-0.8476 if x == "O" else 0.4238 for x in names])
    masses = np.array([ATOMIC_MASSES[x] for x in names])
    vel = maxwell_boltzmann(masses, temperature, rng)
    s = System.from_arrays(names, pos, UnitCell.cubic(L), charges=charges, masses=masses, velocities=vel)
    for i, j in bonds:
        s.add_bond(i, j)
    lj = PairInteraction(LennardJones(sigma=3.166, epsilon=units.from_(0.650, "kJ/mol")), cutoff)
    lj.set_restriction(PairRestriction.INTER_MOLECULAR)
    s.set_pair_potential(("O", "O"), lj)
    s.set_pair_potential(("O", "H"), PairInteraction(NullPotential(), cutoff))
    s.set_pair_potential(("H", "H"), PairInteraction(NullPotential(), cutoff))
    s.set_bond_potential(("O", "H"), Harmonic(k=units.from_(1054.2, "kcal/mol/A^2"), x0=1.0))
    s.set_angle_potential(("H", "O", "H"), Harmonic(k=units.from_(75.9, "kcal/mol"), x0=theta))
    ew = Ewald(cutoff, kmax)
    ew.set_restriction(PairRestriction.INTER_MOLECULAR)
    s.set_coulomb_potential(ew)
    return s


def lj_fluid(natoms, seed=1, density=500 / 28.7 ** 3, temperature=120.0, cutoff=8.5, jitter=0.1):
    """Argon fluid of any size at the density of the 500-atom FCC system:
    `natoms` atoms on randomly chosen sites of a simple cubic lattice with
    U(+-jitter) A noise, shifted LJ (sigma 3.405 A, eps 0.997 kJ/mol) cut at
    min(cutoff, L/2 - 0.01), Maxwell-Boltzmann velocities."""
    rng = np.random.default_rng(seed)
    L = (natoms / density) ** (1.0 / 3.0)
    m = int(math.ceil(natoms ** (1.0 / 3.0) - 1e-9))
    sites = np.array([[i, j, k] for i in range(m) for j in range(m) for k in range(m)], float)
    sites = sites[np.sort(rng.permutation(len(sites))[:natoms])]
    pos = (sites + 0.5) * (L / m) + rng.uniform(-jitter, jitter, size=(natoms, 3))
    masses = np.full(natoms, 39.948)
    vel = maxwell_boltzmann(masses, temperature, rng)
    s = System.from_arrays(["Ar"] * natoms, pos, UnitCell.cubic(L), masses=masses, velocities=vel)
    rc = min(cutoff, 0.5 * L - 0.01)
    s.set_pair_potential(("Ar", "Ar"), PairInteraction.shifted(LennardJones(sigma=LJ_SIGMA, epsilon=LJ_EPSILON), rc))
    return s


def kob_andersen(natoms=1_000_000, seed=6, rho_star=1.2, t_star=1.0, rc_sigma=2.5, jitter=0.05):
    """Kob-Andersen binary mixture (W. Kob, H. C. Andersen, PRE 51, 4626
    (1995)): 80 % A, 20 % B; eps_AB = 1.5, eps_BB = 0.5, sigma_AB = 0.8,
    sigma_BB = 0.88 (units of A-A), all masses 39.948.  Atoms on randomly
    chosen sites of a simple cubic lattice at rho*, jittered by
    U(+-jitter sigma), kinds assigned at random (seeded); Maxwell-Boltzmann
    velocities at T*; plain LJ cut at rc_sigma sigma_AA for every pair."""
    rng = np.random.default_rng(seed)
    sigma, eps = LJ_SIGMA, LJ_EPSILON
    L = (natoms / rho_star) ** (1.0 / 3.0) * sigma
    m = int(math.ceil(natoms ** (1.0 / 3.0) - 1e-9))
    pick = np.sort(rng.permutation(m ** 3)[:natoms])
    sites = np.stack([pick // (m * m), (pick // m) % m, pick % m], axis=1).astype(float)
    pos = (sites + 0.5) * (L / m) + rng.uniform(-jitter, jitter, size=(natoms, 3)) * sigma
    kinds = np.where(rng.permutation(natoms) < int(0.8 * natoms), "A", "B").tolist()
    masses = np.full(natoms, 39.948)
    temperature = t_star * eps / K_BOLTZMANN
    vel = maxwell_boltzmann(masses, temperature, rng)
    s = System.from_arrays(kinds, pos, UnitCell.cubic(L), masses=masses, velocities=vel)
    rc = rc_sigma * sigma
    for (a, b), (e, sg) in {("A", "A"): (1.0, 1.0), ("A", "B"): (1.5, 0.8), ("B", "B"): (0.5, 0.88)}.items():
        s.set_pair_potential((a, b), PairInteraction(LennardJones(sigma=sg * sigma, epsilon=e * eps), rc))
    return s


def water_box(nmol, seed=3, cutoff=10.0):
    """SPC/E water of `nmol` molecules at 1.0 g/cm^3 (0.03343 molecules/A^3),
    cutoff min(cutoff, L/2 - 0.01), kmax scaled with the box (10 at 44.8 A,
    at least 5)."""
    L = (nmol / 0.033428) ** (1.0 / 3.0)
    rc = min(cutoff, 0.5 * L - 0.01)
    kmax = max(5, int(round(10 * L / 44.8)))
    return spce(nmol=nmol, L=L, seed=seed, cutoff=rc, kmax=kmax)


def md_nacl_nvt(system, dt=1.0, temperature=1200.0, tau=100.0):
    md = MolecularDynamics(VelocityVerlet(dt))
    md.set_thermostat(BerendsenThermostat(temperature, tau))
    return Simulation(md)


def md_nve(dt):
    return Simulation(MolecularDynamics(VelocityVerlet(dt)))



def lj_mc_chain(natoms, seed=4, rho_star=0.8, t_star=1.2, rc_sigma=2.5, delta_sigma=0.3, jitter=0.05):
    """LJ fluid for single-particle Monte Carlo: `natoms` argon atoms on
    randomly chosen sites of a simple cubic lattice at density rho*, jittered
    by U(+-jitter sigma), plain LJ cut at rc_sigma sigma.  Returns
    (system, beta, delta) with beta = 1 / (kB T* eps / kB) and delta in A."""
    rng = np.random.default_rng(seed)
    sigma = LJ_SIGMA
    L = (natoms / rho_star) ** (1.0 / 3.0) * sigma
    m = int(math.ceil(natoms ** (1.0 / 3.0) - 1e-9))
    sites = np.array([[i, j, k] for i in range(m) for j in range(m) for k in range(m)], float)
    sites = sites[np.sort(rng.permutation(len(sites))[:natoms])]
    pos = (sites + 0.5) * (L / m) + rng.uniform(-jitter, jitter, size=(natoms, 3)) * sigma
    s = System.from_arrays(["Ar"] * natoms, pos, UnitCell.cubic(L), masses=np.full(natoms, 39.948))
    s.set_pair_potential(("Ar", "Ar"), PairInteraction(LennardJones(sigma=sigma, epsilon=LJ_EPSILON),
                                                       rc_sigma * sigma))
    temperature = t_star * LJ_EPSILON / K_BOLTZMANN
    beta = 1.0 / (K_BOLTZMANN * temperature)
    return s, beta, delta_sigma * sigma


def mc_translate(temperature, delta, seed=0):
    mc = MonteCarlo(temperature, seed=seed)
    mc.add(Translate(delta))
    return Simulation(mc)


def serving_stream(md_jobs=256, steps_scale=0.02, seed=5, precision=None, fp64_every=0, rng=None):
    """The MD requests of the mixed serving stream: `md_jobs` jobs, each LJ
    argon of U{64..8192} atoms (dt 1 fs) or SPC/E water of U{216..12000}
    molecules (dt 0.5 fs) with equal odds, running U{1000..20000} steps
    scaled by `steps_scale`.  Returns (name, system, steps, md options)
    tuples in submission order; every `fp64_every`-th job (0: none) runs the
    FP64 form, the others `precision`.  Draws from `rng` when given (a
    caller that goes on drawing from the same stream), else from a
    generator seeded with `seed`."""
    rng = np.random.default_rng(seed) if rng is None else rng
    out = []
    for k in range(md_jobs):
        steps = max(1, int(rng.integers(1000, 20001) * steps_scale))
        form = precision
        if fp64_every and k % fp64_every == 0:
            form = 0
        if rng.random() < 0.5:
            n = int(rng.integers(64, 8193))
            opts = {"dt": 1.0}
            if form is not None:
                opts["precision"] = form
            out.append((f"lj-{n}", lj_fluid(n, seed=100 + k), steps, opts))
        else:
            nmol = int(rng.integers(216, 12001))
            opts = {"dt": 0.5}
            if form is not None:
                opts["precision"] = form
            out.append((f"water-{nmol}", water_box(nmol, seed=100 + k), steps, opts))
    return out
