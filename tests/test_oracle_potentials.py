"""Pins the oracle's other potential kinds and TableComputation
(oracle/lumol_oracle.cpp, Potential) to the reference's own known answers
(lumol-core/src/energy/functions.rs tests and doc tests,
lumol-core/src/energy/computations.rs tests) and to what the mathematics
fixes: force = -dV/dr, tail_energy = integral of r^2 V from rc, tail_virial =
integral of r^3 f from rc (with the terms by which the reference's
Buckingham, Born-Mayer-Huggins and Gaussian tails differ from the
integrals, and its Morse force from -dV/dr, reproduced as they are), exact table
values at the grid points, linear interpolation between them, and, for a
whole system, F = -dE/dx.

CPU only; no GPU involved.
"""
import math

import numpy as np
import pytest
from scipy import integrate

import lumol_oracle as lo
import reference_values as R
from helpers import other_potentials_system
from lumol_amd._native import (LG_POT_BORN, LG_POT_BUCKINGHAM, LG_POT_COSINE_HARMONIC, LG_POT_GAUSSIAN,
                               LG_POT_HARMONIC, LG_POT_LJ, LG_POT_MIE, LG_POT_MORSE, LG_POT_TORSION)
from test_oracle_golden import _fd_force, ulps_eq

# the C ABI's parameter order of each kind
KINDS = {
    "buckingham": (LG_POT_BUCKINGHAM, ("a", "c", "rho")),
    "born": (LG_POT_BORN, ("a", "c", "d", "sigma", "rho")),
    "morse": (LG_POT_MORSE, ("a", "x0", "depth")),
    "gaussian": (LG_POT_GAUSSIAN, ("a", "b")),
    "mie": (LG_POT_MIE, ("sigma", "epsilon", "n", "m")),
    "cosine_harmonic": (LG_POT_COSINE_HARMONIC, ("k", "x0")),
    "torsion": (LG_POT_TORSION, ("k", "delta", "n")),
}


def kind_and_params(name, fields):
    base = name.replace("_doc", "")
    kind, order = KINDS[base]
    return kind, [float(fields[k]) for k in order]


def value(kind, params, what, x, table=None):
    return lo.potential_value(kind, params, what, x, table)


@pytest.mark.parametrize("name", sorted(R.OTHER_POTENTIALS))
def test_known_answers(name):
    fields, checks = R.OTHER_POTENTIALS[name]
    kind, params = kind_and_params(name, fields)
    for what, x, expected, check in checks:
        got = value(kind, params, what, x)
        if check == "eq":
            assert got == expected, (what, x, got, expected)
        elif check == "ulps":
            assert ulps_eq(got, expected), (what, x, got, expected)
        else:
            mode, eps = check
            if mode == "abs":
                assert abs(got - expected) <= eps, (what, x, got, expected)
            else:  # approx::relative_eq!: absolute epsilon, then f64::EPSILON relative
                assert abs(got - expected) <= eps or \
                    abs(got - expected) <= 2.220446049250313e-16 * max(abs(got), abs(expected)), (what, x, got)


# (kind, params, points) away from singularities
SMOOTH = [
    (LG_POT_BUCKINGHAM, [2.0, 1.0, 2.0], [1.5, 2.0, 4.0, 7.0]),
    (LG_POT_BORN, [2.0, 1.0, 0.5, 2.0, 2.0], [1.5, 2.0, 4.0, 7.0]),
    (LG_POT_GAUSSIAN, [8.0, 2.0], [0.1, 0.5, 1.0, 1.7]),
    (LG_POT_MIE, [2.0, 0.8, 14.0, 7.0], [1.9, 2.5, 3.3, 6.0]),
    (LG_POT_MIE, [3.1, 1.3, 10.5, 5.5], [2.9, 3.5, 5.0]),
    (LG_POT_COSINE_HARMONIC, [50.0, 2.0], [0.5, 1.9, 2.3, 3.0]),
    (LG_POT_TORSION, [5.0, 3.0, 3.0], [-2.0, 0.4, 1.1, 4.0]),
]


@pytest.mark.parametrize("kind,params,xs", SMOOTH)
def test_force_is_minus_derivative(kind, params, xs):
    # the reference's recipe (functions.rs tests: (e0 - e1) / EPS vs force),
    # here as a central difference
    h = 1e-6
    for x in xs:
        f = value(kind, params, "force", x)
        fd = -(value(kind, params, "energy", x + h) - value(kind, params, "energy", x - h)) / (2 * h)
        assert abs(fd - f) <= 1e-7 * max(1.0, abs(f)), (kind, x, fd, f)


def test_morse_force_as_the_reference_has_it():
    # functions.rs:420-423 returns 2 depth (1 - e^2) a with e = exp(a (x0 - r)),
    # which is -dV/dr only where e = 1 (r = x0, the point its own test checks,
    # :754-757); elsewhere -dV/dr = 2 depth a e (e - 1).  Kept as it is.
    a, x0, depth = 2.0, 1.3, 4.0
    h = 1e-6
    for x in (0.9, 1.3, 2.0, 3.5):
        e = math.exp(a * (x0 - x))
        f = value(LG_POT_MORSE, [a, x0, depth], "force", x)
        fd = -(value(LG_POT_MORSE, [a, x0, depth], "energy", x + h) -
               value(LG_POT_MORSE, [a, x0, depth], "energy", x - h)) / (2 * h)
        assert abs(fd - 2 * depth * a * e * (e - 1)) <= 1e-7 * max(1.0, abs(fd)), x
        assert abs(f - 2 * depth * (1 - e * e) * a) <= 1e-14 * max(1.0, abs(f)), x
        if x == x0:
            assert abs(f - fd) <= 1e-7


def _quad(fn, rc):
    v, err = integrate.quad(fn, rc, np.inf, epsabs=0.0, epsrel=1e-12, limit=400)
    return v


TAILS = [
    (LG_POT_LJ, [2.0, 0.8], [2.5, 5.0, 14.42]),
    (LG_POT_BUCKINGHAM, [2.0, 1.0, 2.0], [3.0, 10.0]),
    (LG_POT_BORN, [2.0, 1.0, 0.5, 2.0, 2.0], [3.0, 10.0]),
    (LG_POT_GAUSSIAN, [8.0, 2.0], [0.8, 1.5, 2.5]),
    (LG_POT_MIE, [2.0, 0.8, 14.0, 7.0], [3.0, 6.0]),
    (LG_POT_MIE, [3.1, 1.3, 10.5, 5.5], [4.0, 8.0]),
]


@pytest.mark.parametrize("kind,params,rcs", TAILS)
def test_tails_are_integrals(kind, params, rcs):
    for rc in rcs:
        e_int = _quad(lambda r: r * r * value(kind, params, "energy", r), rc)
        w_int = _quad(lambda r: r ** 3 * value(kind, params, "force", r), rc)
        te, tw = value(kind, params, "tail_energy", rc), value(kind, params, "tail_virial", rc)
        if kind in (LG_POT_BUCKINGHAM, LG_POT_BORN):
            # The reference's closed forms differ from the integrals, and are
            # kept as they are: the tail energy's factor is
            # rc^2 - 2 rc rho + 2 rho^2 where the integral of r^2 A e^(-r/rho)
            # gives rc^2 + 2 rc rho + 2 rho^2 (functions.rs:307, :374); the
            # C/r^6 term enters the tail virial as -20 C / rc^3 where the
            # integral gives -2 C / rc^3, and Buckingham adds a constant 8.0
            # (:317, :392).
            a, c = params[0], params[1]
            if kind == LG_POT_BUCKINGHAM:
                rho, ex = params[2], math.exp(-rc / params[2])
            else:
                rho, ex = params[4], math.exp((params[3] - rc) / params[4])
            e_int = e_int - a * rho * ex * 4.0 * rc * rho
            w_int = w_int + 2.0 * c / rc ** 3 - 20.0 * c / rc ** 3 + (8.0 if kind == LG_POT_BUCKINGHAM else 0.0)
        if kind == LG_POT_GAUSSIAN:
            # functions.rs:493-498: the Gaussian tail virial is the integral
            # with the opposite sign (LJ, Mie: the integral itself); kept as it is
            w_int = -w_int
        assert abs(te - e_int) <= 1e-9 * max(abs(e_int), 1e-12), (kind, rc, te, e_int)
        assert abs(tw - w_int) <= 1e-9 * max(abs(w_int), 1e-12), (kind, rc, tw, w_int)


def test_mie_tail_vanishes_for_m_at_most_3():
    # functions.rs:569-571, 582-584
    assert value(LG_POT_MIE, [2.0, 1.0, 6.0, 3.0], "tail_energy", 5.0) == 0.0
    assert value(LG_POT_MIE, [2.0, 1.0, 6.0, 2.5], "tail_virial", 5.0) == 0.0


def test_invalid_parameters_are_refused():
    # Gaussian::new asserts b > 0 (functions.rs:465); Mie::new asserts m < n (:544)
    assert math.isnan(value(LG_POT_GAUSSIAN, [8.0, -2.0], "energy", 1.0))
    assert math.isnan(value(LG_POT_GAUSSIAN, [8.0, 0.0], "energy", 1.0))
    assert math.isnan(value(LG_POT_MIE, [2.0, 1.0, 6.0, 6.0], "energy", 3.0))
    assert math.isnan(value(LG_POT_MIE, [2.0, 1.0, 6.0, 12.0], "energy", 3.0))


# --- TableComputation (computations.rs) --------------------------------------------------
def test_table_known_answers():
    t = R.TABLE_HARMONIC
    table = (t["size"], t["max"])
    for x, e, f in t["values"]:
        assert value(LG_POT_HARMONIC, [t["k"], t["x0"]], "energy", x, table) == e, x
        assert value(LG_POT_HARMONIC, [t["k"], t["x0"]], "force", x, table) == f, x
    d = R.TABLE_DOC
    for x, e, _ in d["values"]:
        assert value(LG_POT_HARMONIC, [d["k"], d["x0"]], "energy", x, (d["size"], d["max"])) == e
    # computations.rs:203-210: the tails are the wrapped potential's
    for what in ("tail_energy", "tail_virial"):
        assert value(LG_POT_LJ, [2.0, 50.0], what, 5.0, (1000, 4.0)) == value(LG_POT_LJ, [2.0, 50.0], what, 5.0)


def test_table_is_linear_interpolation():
    size, mx = 500, 6.0
    delta = mx / size
    params = [2.0, 0.8]
    for i in (150, 200, 333, 498):
        r = i * delta
        for what in ("energy", "force"):
            direct = value(LG_POT_LJ, params, what, r)
            assert value(LG_POT_LJ, params, what, r, (size, mx)) == direct
            nxt = value(LG_POT_LJ, params, what, (i + 1) * delta)
            mid = value(LG_POT_LJ, params, what, r + 0.5 * delta, (size, mx))
            assert abs(mid - 0.5 * (direct + nxt)) <= 1e-12 * max(abs(direct), abs(nxt)), (i, what)
    # from the last bin on the table gives 0 (bin < size - 1 fails)
    assert value(LG_POT_LJ, params, "energy", (size - 1) * delta, (size, mx)) == 0.0
    assert value(LG_POT_LJ, params, "force", mx + 1.0, (size, mx)) == 0.0


# --- whole systems ---------------------------------------------------------------------------
@pytest.mark.parametrize("coulomb", ["wolf", None])
def test_other_potentials_forces_are_minus_energy_gradient(coulomb):
    from lumol_amd import Harmonic, units
    s = other_potentials_system(variant=1, coulomb=coulomb, n_mol=8, L=15.5)
    # the reference's Morse force is not -dV/dr (test above): harmonic bonds here
    for key in list(s.bond_potentials):
        s.set_bond_potential(key, Harmonic(k=units.from_(2000.0, "kJ/mol/A^2"), x0=1.54))
    o = lo.OracleSystem(s)
    f = o.forces()
    scale = max(1.0, np.abs(f).max())
    for atom in (0, 1, 2, 3, 9, 18):
        for axis in range(3):
            fd = _fd_force(o, s, atom, axis)
            assert abs(fd - f[atom, axis]) <= 2e-6 * scale, (atom, axis, fd, f[atom, axis])
    assert np.abs(f.sum(0)).max() < 1e-9 * scale * len(f)
    t = o.energy_terms()
    assert t["bonds"] != 0.0 and t["angles"] != 0.0 and t["dihedrals"] != 0.0 and t["pairs_tail"] != 0.0


def test_table_system_close_to_direct():
    # variant 2 tabulates the CB-CB Lennard-Jones on 2000 points: within the
    # interpolation error of the same system evaluated directly
    from lumol_amd import LennardJones, PairInteraction, units
    s = other_potentials_system(variant=2, coulomb=None, n_mol=8, L=15.5)
    o = lo.OracleSystem(s)
    f, e = o.forces(), o.energy_terms()
    p = s.pairs[("CB", "CB")]
    direct = PairInteraction(LennardJones(3.95, units.from_(0.382, "kJ/mol")), p.cutoff())
    direct.enable_tail_corrections()
    direct.set_restriction(p.restriction_)
    s.set_pair_potential(("CB", "CB"), direct)
    o2 = lo.OracleSystem(s)
    f2, e2 = o2.forces(), o2.energy_terms()
    assert np.abs(f - f2).max() <= 1e-4 * np.abs(f2).max()
    assert abs(e["pairs"] - e2["pairs"]) <= 1e-5 * abs(e2["pairs"])
    assert e["pairs_tail"] == e2["pairs_tail"]
    assert e["pairs"] != e2["pairs"]


# --- MolecularVirial (compute.rs:278-364, ewald.rs:533-570, 736-753, wolf.rs:285-324) ----
def _com_displacements(s):
    """r_i - centre of mass of i's molecule, per atom, and the strained
    positions of x -> (1 + eps) x applied to the centres of mass only."""
    pos = s.positions
    m = s.masses
    d = np.zeros_like(pos)
    com = np.zeros_like(pos)
    for a, b in s.molecules():
        c = (m[a:b, None] * pos[a:b]).sum(0) / m[a:b].sum()
        d[a:b] = pos[a:b] - c
        com[a:b] = c
    return d, com


@pytest.mark.parametrize("coulomb", ["ewald", None])
def test_molecular_virial_trace_is_minus_rigid_strain_derivative(coulomb):
    # tr W = -dU/d(eps) when the centres of mass follow an isotropic strain
    # and each molecule moves rigidly (compute.rs:258-276, Allen & Tildesley
    # eq. 2.63).  Only the trace: the reference's tensor form
    # w_ab (r_ab . r_ij) / r_ab^2 has the trace of the exact r_ab (x) r_ij form.
    # (Wolf's pair force carries the force_constant shift, wolf.rs:107-117,
    # so it is not -dU/dr of the Wolf energy: the identity test below covers it.)
    from helpers import random_molecular_system
    from lumol_amd import UnitCell
    s = random_molecular_system(n_mol=10, L=13.5, coulomb=coulomb)
    w = lo.OracleSystem(s).molecular_virial()
    _, com = _com_displacements(s)
    L0 = s.cell.lengths[0]
    h = 1e-6
    es = []
    for sign in (1, -1):
        s2 = random_molecular_system(n_mol=10, L=13.5, coulomb=coulomb)
        s2.cell = UnitCell.cubic(L0 * (1 + sign * h))
        s2.positions = s.positions + sign * h * com
        es.append(lo.OracleSystem(s2).potential_energy())
    fd = -(es[0] - es[1]) / (2 * h)
    assert abs(fd - np.trace(w)) <= 1e-6 * abs(fd), (fd, np.trace(w))


@pytest.mark.parametrize("coulomb", ["ewald", "wolf", None])
def test_molecular_virial_trace_identity(coulomb):
    # tr W_molecular = tr W_atomic - sum_i F_i . (r_i - com_i): intra-molecular
    # terms cancel between the two (a bond or excluded pair gives
    # f . r_ab on both sides; angle forces have sum_i F_i . r_i = 0), and
    # the inter-molecular pairs turn r_ab into r_ij
    from helpers import random_molecular_system
    s = random_molecular_system(n_mol=12, L=14.0, coulomb=coulomb)
    o = lo.OracleSystem(s)
    d, _ = _com_displacements(s)
    f = o.forces()
    expected = np.trace(o.virial()) - (f * d).sum()
    got = np.trace(o.molecular_virial())
    assert abs(got - expected) <= 1e-10 * np.abs(o.virial()).max(), (got, expected)


def test_molecular_virial_of_atoms_is_atomic_virial():
    # one-atom molecules: r_ab = r_ij, so every weight is 1, and the k-space
    # correction vanishes (r_i - com = 0)
    from lumol_amd.workloads import nacl
    s = nacl(ncell=3, cutoff=8.0, kmax=6)
    o = lo.OracleSystem(s, nthreads=4)
    wa, wm = o.virial(), o.molecular_virial()
    assert np.abs(wa - wm).max() <= 1e-12 * np.abs(wa).max()
