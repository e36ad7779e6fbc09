"""ctypes wrapper of liblumol_oracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  It pushes the same system description a
``lumol_amd.System`` holds into the CPU FP64 restatement of the reference
(oracle/lumol_oracle.cpp) so results can be compared on identical inputs.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblumol_oracle.so")

_lib = None
D = ctypes.c_double
DP = ctypes.POINTER(ctypes.c_double)
U32 = ctypes.c_uint32
U32P = ctypes.POINTER(ctypes.c_uint32)
U64 = ctypes.c_uint64
SZ = ctypes.c_size_t
I = ctypes.c_int
P = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run oracle/build.py")
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "lo_last_error": (ctypes.c_char_p, []),
            "lo_create": (P, [SZ]),
            "lo_destroy": (None, [P]),
            "lo_set_threads": (I, [P, I]),
            "lo_set_cell": (I, [P, D, D, D]),
            "lo_set_positions": (I, [P, DP]),
            "lo_get_positions": (I, [P, DP]),
            "lo_set_velocities": (I, [P, DP]),
            "lo_get_velocities": (I, [P, DP]),
            "lo_set_masses": (I, [P, DP]),
            "lo_set_charges": (I, [P, DP]),
            "lo_set_kinds": (I, [P, U32P, U32]),
            "lo_set_bonds": (I, [P, SZ, U32P]),
            "lo_set_pair": (I, [P, U32, U32, I, DP, D, I, I, I, D]),
            "lo_set_pair_table": (I, [P, U32, U32, I, DP, U64, D, D, I, I, I, D]),
            "lo_potential_value": (D, [I, DP, U64, D, I, D]),
            "lo_set_bond_potential": (I, [P, U32, U32, I, DP]),
            "lo_set_angle_potential": (I, [P, U32, U32, U32, I, DP]),
            "lo_set_dihedral_potential": (I, [P, U32, U32, U32, U32, I, DP]),
            "lo_set_ewald": (I, [P, D, I, D, I, D]),
            "lo_set_wolf": (I, [P, D, I, D]),
            "lo_forces": (I, [P, DP]),
            "lo_energy": (I, [P, DP]),
            "lo_ewald_components": (I, [P, DP, DP, DP]),
            "lo_ewald_kvectors": (I, [P, ctypes.POINTER(I)]),
            "lo_virial": (I, [P, DP]),
            "lo_molecular_virial": (I, [P, DP]),
            "lo_kinetic_energy": (D, [P]),
            "lo_temperature": (D, [P]),
            "lo_pair_counts": (I, [P, ctypes.POINTER(U64), ctypes.POINTER(U64)]),
            "lo_pairs_within": (U64, [P, D, U32P, U64]),
            "lo_md_reset": (I, [P]),
            "lo_md_run": (I, [P, I, D, U64, D, D]),
            "lo_cache_init": (I, [P]),
            "lo_cache_energy": (D, [P]),
            "lo_move_cost": (I, [P, U32, DP, DP]),
            "lo_move_commit": (I, [P, I]),
            "lo_move_discard": (I, [P]),
            "lo_mc_run": (I, [P, U64, U64, U64, D, D, ctypes.POINTER(U64)]),
            "lo_philox4x32_10": (None, [U32P, U32P, U32P]),
            "lo_lj_energy": (D, [D, D, D]),
            "lo_lj_force": (D, [D, D, D]),
            "lo_lj_tail_energy": (D, [D, D, D]),
            "lo_lj_tail_virial": (D, [D, D, D]),
            "lo_pair_interaction_energy": (D, [I, D, D, D, I, D]),
            "lo_pair_interaction_force": (D, [I, D, D, D, D]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _chk(status):
    if status != 0:
        raise RuntimeError("oracle: " + lib().lo_last_error().decode())


def _padded(params):
    pr = np.zeros(8)
    pr[:len(params)] = params
    return pr


def _dp(a):
    return a.ctypes.data_as(DP)


class OracleSystem:
    """Oracle instance holding the same spec as a lumol_amd.System."""

    def __init__(self, system, nthreads=1):
        L = lib()
        self.n = system.size()
        self.h = L.lo_create(self.n)
        _chk(L.lo_set_threads(self.h, nthreads))
        a, b, c = system.cell.lengths
        _chk(L.lo_set_cell(self.h, a, b, c))
        kinds = np.array([system.kinds[x] for x in system.names], dtype=np.uint32)
        _chk(L.lo_set_kinds(self.h, kinds.ctypes.data_as(U32P), max(1, len(system.kinds))))
        pos = np.ascontiguousarray(system.positions, dtype=float)
        vel = np.ascontiguousarray(system.velocities, dtype=float)
        _chk(L.lo_set_positions(self.h, _dp(pos)))
        _chk(L.lo_set_velocities(self.h, _dp(vel)))
        q = np.ascontiguousarray(system.charges, dtype=float)
        _chk(L.lo_set_charges(self.h, _dp(q)))
        m = np.ascontiguousarray(system.masses, dtype=float)
        _chk(L.lo_set_masses(self.h, _dp(m)))
        bonds = np.ascontiguousarray(np.array(system.bonds, dtype=np.uint32).reshape(-1, 2))
        _chk(L.lo_set_bonds(self.h, len(bonds), bonds.ctypes.data_as(U32P)))
        for (na, nb), inter in system.pairs.items():
            pot = inter.potential
            table = getattr(pot, "table", None)  # TableComputation: (size, max)
            if table is not None:
                pot = pot.potential
            pr = _padded(pot.params())
            if table is not None:
                _chk(L.lo_set_pair_table(self.h, system.kinds[na], system.kinds[nb], pot.kind, _dp(pr),
                                         int(table[0]), float(table[1]), inter.cutoff(), int(inter.shifted_),
                                         int(inter.tail), int(inter.restriction_), float(inter.scale14)))
            else:
                _chk(L.lo_set_pair(self.h, system.kinds[na], system.kinds[nb], pot.kind, _dp(pr),
                                   inter.cutoff(), int(inter.shifted_), int(inter.tail), int(inter.restriction_),
                                   float(inter.scale14)))
        for names, pot in system.bond_potentials.items():
            pr = _padded(pot.params())
            _chk(L.lo_set_bond_potential(self.h, system.kinds[names[0]], system.kinds[names[1]], pot.kind,
                                         _dp(pr)))
        for names, pot in system.angle_potentials.items():
            pr = _padded(pot.params())
            k = [system.kinds[x] for x in names]
            _chk(L.lo_set_angle_potential(self.h, k[0], k[1], k[2], pot.kind, _dp(pr)))
        for names, pot in system.dihedral_potentials.items():
            pr = _padded(pot.params())
            k = [system.kinds[x] for x in names]
            _chk(L.lo_set_dihedral_potential(self.h, k[0], k[1], k[2], k[3], pot.kind, _dp(pr)))
        coul = system.coulomb
        if coul is not None and type(coul).__name__ == "Ewald":
            _chk(L.lo_set_ewald(self.h, coul.cutoff, coul.kmax, coul.alpha, coul.restriction, 1.0))
        elif coul is not None and type(coul).__name__ == "Wolf":
            _chk(L.lo_set_wolf(self.h, coul.cutoff, coul.restriction, coul.scale14))

    def __del__(self):
        try:
            lib().lo_destroy(self.h)
        except Exception:
            pass

    def set_threads(self, n):
        _chk(lib().lo_set_threads(self.h, n))

    def forces(self):
        out = np.zeros((self.n, 3))
        _chk(lib().lo_forces(self.h, _dp(out)))
        return out

    def energy_terms(self):
        out = np.zeros(8)
        _chk(lib().lo_energy(self.h, _dp(out)))
        keys = ["pairs", "pairs_tail", "bonds", "angles", "dihedrals", "coulomb", "global", "total"]
        return dict(zip(keys, out))

    def potential_energy(self):
        return self.energy_terms()["total"]

    def virial(self):
        out = np.zeros(9)
        _chk(lib().lo_virial(self.h, _dp(out)))
        return out.reshape(3, 3)

    def molecular_virial(self):
        out = np.zeros(9)
        _chk(lib().lo_molecular_virial(self.h, _dp(out)))
        return out.reshape(3, 3)

    def ewald_components(self):
        e = np.zeros(3)
        wr = np.zeros(9)
        wk = np.zeros(9)
        _chk(lib().lo_ewald_components(self.h, _dp(e), _dp(wr), _dp(wk)))
        return {"real": e[0], "self": e[1], "kspace": e[2], "real_virial": wr.reshape(3, 3),
                "kspace_virial": wk.reshape(3, 3)}

    def ewald_kvectors(self):
        c = I()
        _chk(lib().lo_ewald_kvectors(self.h, ctypes.byref(c)))
        return c.value

    def pair_counts(self):
        a, b = U64(), U64()
        _chk(lib().lo_pair_counts(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def pairs_within(self, rc):
        n = lib().lo_pairs_within(self.h, rc, None, 0)
        out = np.zeros((n, 2), dtype=np.uint32)
        lib().lo_pairs_within(self.h, rc, out.ctypes.data_as(U32P), n)
        return out

    def kinetic_energy(self):
        return lib().lo_kinetic_energy(self.h)

    def temperature(self):
        return lib().lo_temperature(self.h)

    def positions(self):
        out = np.zeros((self.n, 3))
        _chk(lib().lo_get_positions(self.h, _dp(out)))
        return out

    def velocities(self):
        out = np.zeros((self.n, 3))
        _chk(lib().lo_get_velocities(self.h, _dp(out)))
        return out

    def set_positions(self, pos):
        pos = np.ascontiguousarray(pos, dtype=float)
        _chk(lib().lo_set_positions(self.h, _dp(pos)))

    def set_velocities(self, vel):
        vel = np.ascontiguousarray(vel, dtype=float)
        _chk(lib().lo_set_velocities(self.h, _dp(vel)))

    def set_cell(self, a, b, c):
        _chk(lib().lo_set_cell(self.h, a, b, c))

    def md_reset(self):
        _chk(lib().lo_md_reset(self.h))

    def md_run(self, integrator, dt, nsteps, berendsen_T=-1.0, tau=1.0):
        """berendsen_T >= 0 applies the Berendsen thermostat after each step."""
        _chk(lib().lo_md_run(self.h, integrator, dt, nsteps, berendsen_T, tau))

    def cache_init(self):
        _chk(lib().lo_cache_init(self.h))

    def cache_energy(self):
        return lib().lo_cache_energy(self.h)

    def move_cost(self, molid, newpos):
        newpos = np.ascontiguousarray(newpos, dtype=float)
        c = D()
        _chk(lib().lo_move_cost(self.h, molid, _dp(newpos), ctypes.byref(c)))
        return c.value

    def move_commit(self, wrap=True):
        _chk(lib().lo_move_commit(self.h, int(wrap)))

    def move_discard(self):
        _chk(lib().lo_move_discard(self.h))

    def mc_run(self, seed, attempt0, nattempts, beta, delta):
        acc = U64()
        _chk(lib().lo_mc_run(self.h, seed, attempt0, nattempts, beta, delta, ctypes.byref(acc)))
        return acc.value


def potential_value(kind, params, what, x, table=None):
    """Scalar value of one potential: what = "energy", "force",
    "tail_energy" or "tail_virial"; table = (size, max) evaluates its
    TableComputation instead."""
    pr = np.zeros(8)
    pr[:len(params)] = params
    size, mx = table if table is not None else (0, 0.0)
    code = ["energy", "force", "tail_energy", "tail_virial"].index(what)
    return lib().lo_potential_value(kind, _dp(pr), int(size), float(mx), code, float(x))


def philox4x32_10(ctr, key):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().lo_philox4x32_10(c.ctypes.data_as(U32P), k.ctypes.data_as(U32P), o.ctypes.data_as(U32P))
    return o
