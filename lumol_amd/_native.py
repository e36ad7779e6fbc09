"""ctypes binding of liblumol_gpu.so (the C ABI of include/lumol_gpu.h).

There is no fallback: if the library is missing or the device cannot be
opened, every call raises.  The library is built in-tree by
``lumol_amd.build.build()``.
"""
import ctypes
import os
import threading
import warnings

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblumol_gpu.so")

LG_OK = 0
LG_E_INVALID = 1
LG_E_UNSUPPORTED = 2
LG_E_DEVICE = 3
LG_E_STATE = 4
LG_E_NONFINITE = 5

LG_POT_NULL, LG_POT_LJ, LG_POT_HARMONIC = 0, 1, 2
LG_POT_BUCKINGHAM, LG_POT_BORN, LG_POT_MORSE, LG_POT_GAUSSIAN, LG_POT_MIE = 3, 4, 5, 6, 7
LG_POT_COSINE_HARMONIC, LG_POT_TORSION = 8, 9
LG_PRECISION_FP64, LG_PRECISION_MIXED = 0, 1
LG_INTEGRATOR_VELOCITY_VERLET, LG_INTEGRATOR_LEAP_FROG = 0, 1

# every symbol include/lumol_gpu.h declares
EXPORTS = [
    "lg_last_error", "lg_version", "lg_context_create", "lg_context_destroy",
    "lg_system_create", "lg_system_destroy", "lg_system_size", "lg_system_set_cell",
    "lg_system_set_positions", "lg_system_get_positions", "lg_system_set_velocities",
    "lg_system_get_velocities", "lg_system_set_masses", "lg_system_set_charges",
    "lg_system_set_kinds", "lg_system_set_bonds", "lg_system_molecule_count",
    "lg_system_set_precision", "lg_system_set_tuning", "lg_system_get_tuning", "lg_pair_set", "lg_pair_set_table", "lg_bond_set", "lg_angle_set", "lg_dihedral_set",
    "lg_coulomb_ewald", "lg_coulomb_wolf", "lg_coulomb_clear", "lg_forces",
    "lg_potential_energy", "lg_potential_energy_noalloc", "lg_virial", "lg_molecular_virial", "lg_kinetic_energy", "lg_temperature",
    "lg_ewald_kvector_count", "lg_pairs_within", "lg_md_setup", "lg_md_run", "lg_cache_init", "lg_move_cost",
    "lg_move_commit", "lg_move_discard", "lg_cache_energy", "lg_mc_run", "lg_mc_run_batch", "lg_profile_begin",
    "lg_profile_end",
]

LG_KERNEL_NLIST, LG_KERNEL_PAIRS, LG_KERNEL_MOLECULES = 0, 1, 2
LG_KERNEL_KSPACE_RHO, LG_KERNEL_KSPACE_FORCE, LG_KERNEL_INTEGRATE = 3, 4, 5
KERNEL_CLASSES = ["nlist", "pairs", "molecules", "kspace_rho", "kspace_force", "integrate"]


class Profile(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double * 6), ("launches", ctypes.c_uint64 * 6)]


class Tuning(ctypes.Structure):
    _fields_ = [("pair_lanes", ctypes.c_int), ("kspace_split", ctypes.c_int), ("rho_width", ctypes.c_int),
                ("nlist_build", ctypes.c_int), ("md_graph", ctypes.c_int), ("rho_kind", ctypes.c_int)]


class EnergyTerms(ctypes.Structure):
    _fields_ = [(name, ctypes.c_double) for name in (
        "pairs", "pairs_tail", "bonds", "angles", "dihedrals", "coulomb", "global_",
        "total", "coulomb_real", "coulomb_self", "coulomb_kspace")]


class LumolGpuError(RuntimeError):
    """Error returned by the GPU library (status code and message)."""

    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


_lib = None
_lock = threading.Lock()

P = ctypes.c_void_p
D = ctypes.c_double
DP = ctypes.POINTER(ctypes.c_double)
U32 = ctypes.c_uint32
U32P = ctypes.POINTER(ctypes.c_uint32)
SZ = ctypes.c_size_t
I = ctypes.c_int
U64 = ctypes.c_uint64
U64P = ctypes.POINTER(ctypes.c_uint64)
PP = ctypes.POINTER(ctypes.c_void_p)


def _declare(lib):
    sig = {
        "lg_last_error": (ctypes.c_char_p, []),
        "lg_version": (ctypes.c_char_p, []),
        "lg_context_create": (I, [I, ctypes.POINTER(P)]),
        "lg_context_destroy": (I, [P]),
        "lg_system_create": (I, [P, SZ, ctypes.POINTER(P)]),
        "lg_system_destroy": (I, [P]),
        "lg_system_size": (I, [P, ctypes.POINTER(SZ)]),
        "lg_system_set_cell": (I, [P, D, D, D]),
        "lg_system_set_positions": (I, [P, DP]),
        "lg_system_get_positions": (I, [P, DP]),
        "lg_system_set_velocities": (I, [P, DP]),
        "lg_system_get_velocities": (I, [P, DP]),
        "lg_system_set_masses": (I, [P, DP]),
        "lg_system_set_charges": (I, [P, DP]),
        "lg_system_set_kinds": (I, [P, U32P, U32]),
        "lg_system_set_bonds": (I, [P, SZ, U32P]),
        "lg_system_molecule_count": (I, [P, ctypes.POINTER(SZ)]),
        "lg_system_set_precision": (I, [P, I]),
        "lg_system_set_tuning": (I, [P, ctypes.POINTER(Tuning)]),
        "lg_system_get_tuning": (I, [P, ctypes.POINTER(Tuning)]),
        "lg_pair_set": (I, [P, U32, U32, I, DP, D, I, I, I, D]),
        "lg_pair_set_table": (I, [P, U32, U32, I, DP, U64, D, D, I, I, I, D]),
        "lg_bond_set": (I, [P, U32, U32, I, DP]),
        "lg_angle_set": (I, [P, U32, U32, U32, I, DP]),
        "lg_dihedral_set": (I, [P, U32, U32, U32, U32, I, DP]),
        "lg_coulomb_ewald": (I, [P, D, I, D, I]),
        "lg_coulomb_wolf": (I, [P, D, I, D]),
        "lg_coulomb_clear": (I, [P]),
        "lg_forces": (I, [P, DP, I]),
        "lg_potential_energy": (I, [P, ctypes.POINTER(EnergyTerms)]),
        "lg_potential_energy_noalloc": (I, [P, ctypes.POINTER(EnergyTerms)]),
        "lg_virial": (I, [P, DP]),
        "lg_molecular_virial": (I, [P, DP]),
        "lg_kinetic_energy": (I, [P, DP]),
        "lg_temperature": (I, [P, DP]),
        "lg_ewald_kvector_count": (I, [P, ctypes.POINTER(I)]),
        "lg_pairs_within": (I, [P, D, U32P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
        "lg_md_setup": (I, [P, I, D, D, D]),
        "lg_md_run": (I, [P, ctypes.c_uint64]),
        "lg_cache_init": (I, [P, DP]),
        "lg_move_cost": (I, [P, U32, DP, DP]),
        "lg_move_commit": (I, [P, I]),
        "lg_move_discard": (I, [P]),
        "lg_cache_energy": (I, [P, DP]),
        "lg_mc_run": (I, [P, U64, U64, U64, D, D, U64P]),
        "lg_mc_run_batch": (I, [PP, SZ, U64P, U64P, U64P, DP, DP, U64P]),
        "lg_profile_begin": (I, [P, U32]),
        "lg_profile_end": (I, [P, ctypes.POINTER(Profile)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def _hip_loaded():
    """True when the HIP runtime library is already mapped in this process."""
    try:
        with open("/proc/self/maps") as f:
            return "libamdhip64" in f.read()
    except OSError:
        return False


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with lumol_amd.build.build() "
                    "(python __graft_entry__.py build)")
            # Each library context owns one HIP stream, and contexts run
            # concurrently (batched systems, serving lanes) only on distinct
            # hardware queues: with HIP's default of 4 per process a fifth
            # stream shares a queue and serialises behind another.  A lower
            # setting is raised to 8 (4 SPC/E systems at once: 2.15x -> 2.65x
            # the FP64 single-system rate); it takes effect if the HIP runtime
            # is not initialised yet.
            # LUMOL_GPU_KEEP_HW_QUEUES=1 keeps the caller's setting as it is.
            if os.environ.get("LUMOL_GPU_KEEP_HW_QUEUES", "0") != "1":
                try:
                    queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "0"))
                except ValueError:
                    queues = 0
                if queues < 8:
                    if _hip_loaded():
                        warnings.warn("the HIP runtime was loaded before the lumol_gpu library: "
                                      "GPU_MAX_HW_QUEUES=8 may not take effect, and concurrent "
                                      "contexts may share hardware queues", RuntimeWarning, stacklevel=3)
                    os.environ["GPU_MAX_HW_QUEUES"] = "8"
            handle = ctypes.CDLL(LIB_PATH)
            _declare(handle)
            _lib = handle
        return _lib


def check(status):
    if status != LG_OK:
        msg = lib().lg_last_error().decode("utf-8", "replace")
        raise LumolGpuError(status, msg)


def dptr(array):
    return array.ctypes.data_as(DP)


def u32ptr(array):
    return array.ctypes.data_as(U32P)


_contexts = {}
_ctx_lock = threading.Lock()


def context(device=0, slot=0):
    """The lg_context for (device, slot), created on first use and kept for
    the process.  Each context owns one HIP stream: systems on different
    slots of a device run concurrently."""
    key = (device, slot)
    L = lib()
    with _ctx_lock:
        ctx = _contexts.get(key)
        if ctx is None:
            handle = P()
            check(L.lg_context_create(device, ctypes.byref(handle)))
            _contexts[key] = handle
            ctx = handle
    return ctx