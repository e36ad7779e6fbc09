// lg_api.hip -- host side of the C ABI declared in include/lumol_gpu.h.
//
// Owns the device-resident state of each system and sequences the kernels of
// lg_kernels.hpp on the context's stream.  Host code restates the parts of
// the reference that are set-up rather than hot path:
//   Bonding::rebuild (angles, dihedrals, bond distances)
//       lumol-core/src/sys/config/bonding.rs:78-155
//   Ewald k-vector set and factors  lumol-core/src/energy/global/ewald.rs:130-183, 353-378
//   tail corrections                lumol-core/src/sys/energy.rs:63-79, compute.rs:218-228
//   Ewald / Wolf self terms         ewald.rs:619-626, wolf.rs:100-102, 177-207
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <numeric>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/lumol_gpu.h"
#include "lg_kernels.hpp"
#include "lg_mc.hpp"

using namespace lg;

namespace {

thread_local std::string g_error = "";

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define LG_HIP(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return fail(LG_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LG_TRY(expr)              \
    do {                          \
        int r_ = (expr);          \
        if (r_ != LG_OK) return r_; \
    } while (0)

// Bumped whenever any device buffer is (re)allocated or freed: a captured
// HIP graph holds buffer addresses, so it is replayed only while this and
// the system's configuration epoch are unchanged.
static std::atomic<uint64_t> g_alloc_epoch{0};

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;
    int ensure(size_t count) {
        if (count <= cap && p) return LG_OK;
        g_alloc_epoch.fetch_add(1);
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) return fail(LG_E_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
        cap = std::max<size_t>(count, 1);
        return LG_OK;
    }
    void release() {
        if (p) {
            g_alloc_epoch.fetch_add(1);
            (void)hipFree(p);
        }
        p = nullptr;
        cap = 0;
    }
};

// A potential as the reference's structs hold it (parameters in field
// order; see PairParam / BondedParam), plus an optional TableComputation.
struct PotSpec {
    int pot = LG_POT_NULL;
    double p[5] = {0, 0, 0, 0, 0};
    uint64_t tab_n = 0;  // TableComputation size (0: evaluated directly)
    double tab_max = 0;
};
struct PairSpec {
    PotSpec v;
    double rc;
    int shifted, tail, restriction;
    double scale14;
};
using BondedSpec = PotSpec;

// Host evaluation of the potentials in the reference's operation order,
// without fused multiply-adds (Rust fuses none): the shift V(rc), the
// TableComputation tables and the tail corrections are computed here.
#pragma clang fp contract(off)
// functions.rs:79-89, 135-145, 198-207, 244-256, 286-300, 353-367, 414-424,
// 474-482, 552-566
double host_direct(const PotSpec& q, double r, bool force) {
    const double* p = q.p;
    switch (q.pot) {
        case LG_POT_LJ: {
            double x = p[0] / r;
            double x2 = x * x;
            double x4 = x2 * x2;
            double s6 = x2 * x4;
            return force ? -24.0 * p[1] * (s6 - 2.0 * (s6 * s6)) / r : 4.0 * p[1] * (s6 * s6 - s6);
        }
        case LG_POT_HARMONIC: {
            if (force) return p[0] * (p[1] - r);
            double dx = r - p[1];
            return 0.5 * p[0] * dx * dx;
        }
        case LG_POT_BUCKINGHAM: {
            double r3 = r * r * r;
            double ex = std::exp(-r / p[2]);
            if (force) return p[0] / p[2] * ex - 6.0 * p[1] / (r3 * r3 * r);
            return p[0] * ex - p[1] / (r3 * r3);
        }
        case LG_POT_BORN: {
            double r2 = r * r;
            double ex = std::exp((p[3] - r) / p[4]);
            if (force) {
                double r7 = r2 * r2 * r2 * r;
                return p[0] / p[4] * ex - 6.0 * p[1] / r7 + 8.0 * p[2] / (r7 * r2);
            }
            double r6 = r2 * r2 * r2;
            return p[0] * ex - p[1] / r6 + p[2] / (r6 * r2);
        }
        case LG_POT_MORSE: {
            double ex = std::exp((p[1] - r) * p[0]);
            if (force) return 2.0 * p[2] * (1.0 - ex * ex) * p[0];
            double m = 1.0 - ex;
            return p[2] * m * m;
        }
        case LG_POT_GAUSSIAN: {
            double e = -p[0] * std::exp(-p[1] * r * r);
            return force ? 2.0 * p[1] * r * e : e;
        }
        case LG_POT_MIE: {
            double sr = p[0] / r;
            double rep = std::pow(sr, p[1]), att = std::pow(sr, p[2]);
            return force ? p[3] * (p[1] * rep - p[2] * att) / r : p[3] * (rep - att);
        }
        case LG_POT_COSINE_HARMONIC: {
            double dr = std::cos(r) - p[1];
            return force ? p[0] * dr * std::sin(r) : 0.5 * p[0] * dr * dr;
        }
        case LG_POT_TORSION: {
            double x = p[2] * r - p[1];
            return force ? p[0] * p[2] * std::sin(x) : p[0] * (1.0 + std::cos(x));
        }
    }
    return 0.0;
}
// computations.rs:96-117: entries at r = i * max / size
void host_table(const PotSpec& q, std::vector<double>& e, std::vector<double>& f) {
    const double delta = q.tab_max / double(q.tab_n);
    e.resize(q.tab_n);
    f.resize(q.tab_n);
    for (uint64_t i = 0; i < q.tab_n; i++) {
        double r = double(i) * delta;
        e[i] = host_direct(q, r, false);
        f[i] = host_direct(q, r, true);
    }
}
// computations.rs:122-134 (energy of the table, for the shift)
double host_energy(const PotSpec& q, double r) {
    if (q.tab_n == 0) return host_direct(q, r, false);
    const double delta = q.tab_max / double(q.tab_n);
    const double u = std::floor(r / delta);
    const double bin = u > 0.0 ? u : 0.0;
    if (!(bin < double(q.tab_n - 1))) return 0.0;
    const uint64_t b = uint64_t(bin);
    double e0 = host_direct(q, double(b) * delta, false), e1 = host_direct(q, double(b + 1) * delta, false);
    double dx = r - bin * delta;
    return e0 + dx * ((e1 - e0) / delta);
}
// Tails: functions.rs:92-106 (LJ), :302-327 (Buckingham, including its
// `+ 8.0`), :369-395 (BornMayerHuggins), :484-499 (Gaussian), :568-592
// (Mie); zero for Null, Harmonic and Morse.  A table keeps the tails of the
// potential it wraps (computations.rs:148-175).
double host_tail(const PotSpec& q, double rc, bool virial) {
    const double* p = q.p;
    switch (q.pot) {
        case LG_POT_LJ: {
            double s3 = p[0] * p[0] * p[0], rc3 = rc * rc * rc;
            double s9 = s3 * s3 * s3, rc9 = rc3 * rc3 * rc3;
            if (virial) return 8.0 * p[1] * s3 * (2.0 / 3.0 * s9 / rc9 - s3 / rc3);
            return 4.0 / 3.0 * p[1] * s3 * (1.0 / 3.0 * s9 / rc9 - s3 / rc3);
        }
        case LG_POT_BUCKINGHAM: case LG_POT_BORN: {
            const bool born = q.pot == LG_POT_BORN;
            const double rho = born ? p[4] : p[2];
            double rc2 = rc * rc, rc3 = rc2 * rc;
            double ex = born ? std::exp((p[3] - rc) / rho) : std::exp(-rc / rho);
            if (virial) {
                double factor = rc3 + 3.0 * rc2 * rho + 6.0 * rc * rho * rho + 6.0 * rho * rho * rho;
                double w = p[0] * ex * factor - 20.0 * p[1] / rc3;
                return born ? w + 8.0 * p[2] / (5.0 * rc2 * rc3) : w + 8.0;
            }
            double factor = rc2 - 2.0 * rc * rho + 2.0 * rho * rho;
            double e = p[0] * rho * ex * factor - p[1] / (3.0 * rc3);
            return born ? e + p[2] / (5.0 * rc2 * rc3) : e;
        }
        case LG_POT_GAUSSIAN: {
            double erc = host_direct(q, rc, false);
            if (virial)
                return 3.0 * std::sqrt(kPi) * p[0] * std::erfc(std::sqrt(p[1]) * rc) / (4.0 * std::pow(p[1], 3.0 / 2.0)) -
                       erc * rc * (2.0 * p[1] * rc * rc + 3.0) / (2.0 * p[1]);
            return erc * rc / (2.0 * p[1]) -
                   p[0] * std::sqrt(kPi) * std::erfc(std::sqrt(p[1]) * rc) / (4.0 * std::pow(p[1], 3.0 / 2.0));
        }
        case LG_POT_MIE: {
            if (p[2] <= 3.0) return 0.0;
            double sr = p[0] / rc, n3 = p[1] - 3.0, m3 = p[2] - 3.0;
            double rep = std::pow(sr, n3), att = std::pow(sr, m3);
            if (virial) return p[3] * (p[0] * p[0] * p[0]) * (rep * p[1] / n3 - att * p[2] / m3);
            return p[3] * (p[0] * p[0] * p[0]) * (rep / n3 - att / m3);
        }
    }
    return 0.0;
}
#pragma clang fp contract(on)

double host_tail_energy(const PairSpec& p) { return p.tail ? host_tail(p.v, p.rc, false) : 0.0; }
double host_tail_virial(const PairSpec& p) { return p.tail ? host_tail(p.v, p.rc, true) : 0.0; }

// Number of parameters of each kind at the C ABI, and where each kind may be
// used (the reference's trait impls: PairPotential, BondPotential,
// AnglePotential, DihedralPotential; functions.rs, computations.rs).
enum : int { ROLE_PAIR = 1, ROLE_BOND = 2, ROLE_ANGLE = 4, ROLE_DIHEDRAL = 8 };
struct PotInfo {
    int nparams, roles;
    const char* name;
};
PotInfo pot_info(int pot) {
    switch (pot) {
        case LG_POT_NULL: return {0, ROLE_PAIR | ROLE_BOND | ROLE_ANGLE | ROLE_DIHEDRAL, "NullPotential"};
        case LG_POT_LJ: return {2, ROLE_PAIR, "LennardJones"};
        case LG_POT_HARMONIC: return {2, ROLE_PAIR | ROLE_BOND | ROLE_ANGLE | ROLE_DIHEDRAL, "Harmonic"};
        case LG_POT_BUCKINGHAM: return {3, ROLE_PAIR, "Buckingham"};
        case LG_POT_BORN: return {5, ROLE_PAIR, "BornMayerHuggins"};
        case LG_POT_MORSE: return {3, ROLE_PAIR | ROLE_BOND | ROLE_ANGLE | ROLE_DIHEDRAL, "Morse"};
        case LG_POT_GAUSSIAN: return {2, ROLE_PAIR, "Gaussian"};
        case LG_POT_MIE: return {4, ROLE_PAIR, "Mie"};
        case LG_POT_COSINE_HARMONIC: return {2, ROLE_ANGLE | ROLE_DIHEDRAL, "CosineHarmonic"};
        case LG_POT_TORSION: return {3, ROLE_DIHEDRAL, "Torsion"};
    }
    return {-1, 0, "unknown"};
}

// The C ABI's parameters -> PotSpec, validated as the reference's
// constructors assert: Gaussian::new b > 0 (functions.rs:465), Mie::new
// m < n (:544, prefac as there), Torsion's integer multiplicity, and
// CosineHarmonic::new storing cos(x0) (:189-196).
int make_pot(int pot, const double* params, int role, PotSpec& out) {
    PotInfo info = pot_info(pot);
    if (info.nparams < 0) return fail(LG_E_UNSUPPORTED, "unknown potential kind");
    if (!(info.roles & role)) {
        const char* what = role == ROLE_PAIR ? "a pair" : role == ROLE_BOND ? "a bond"
                         : role == ROLE_ANGLE ? "an angle" : "a dihedral";
        return fail(LG_E_UNSUPPORTED, std::string(info.name) + " is not " + what + " potential");
    }
    if (info.nparams > 0 && !params) return fail(LG_E_INVALID, "missing potential parameters");
    out = PotSpec{};
    out.pot = pot;
    for (int k = 0; k < info.nparams; k++) {
        if (!std::isfinite(params[k])) return fail(LG_E_INVALID, "potential parameters must be finite");
        out.p[k] = params[k];
    }
    if (pot == LG_POT_GAUSSIAN && !(out.p[1] > 0.0))
        return fail(LG_E_INVALID, "\"b\" has to be positive in Gaussian potential");
    if (pot == LG_POT_MIE) {
        double sigma = params[0], eps = params[1], n = params[2], m = params[3];
        if (!(m < n)) return fail(LG_E_INVALID, "The repulsive exponent n has to be larger than the attractive exponent m");
        out.p[0] = sigma;
        out.p[1] = n;
        out.p[2] = m;
        out.p[3] = n / (n - m) * std::pow(n / m, m / (n - m)) * eps;
    }
    if (pot == LG_POT_COSINE_HARMONIC) out.p[1] = std::cos(params[1]);
    if (pot == LG_POT_TORSION && !(params[2] >= 0.0 && params[2] == std::floor(params[2])))
        return fail(LG_E_INVALID, "the multiplicity of a torsion must be a non-negative integer");
    return LG_OK;
}

}  // namespace

// ---- scratch of lg_potential_energy_noalloc --------------------------------------------
// Per-block energy and virial partial sums of the allocation-free energy
// entry point, and their 18 totals.  A file-scope device array: every GPU
// the library runs on holds its own copy until the process ends.  Host code
// never uses the name itself (a host-side stand-in): kernels get the device
// address that hipGetSymbolAddress returns for the current device, looked
// up once per device when a context is created.
constexpr size_t kNoallocScratchDoubles = size_t(1) << 20;  // 8 MiB
static __device__ __align__(256) double scratch[kNoallocScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current).
int scratch_of_device(int device, double** out) {
    if (device < 0 || device >= kMaxDevices) return fail(LG_E_UNSUPPORTED, "device index above the supported range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(LG_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[device];
    return LG_OK;
}
}  // namespace

struct lg_context {
    int device = 0;
    double* scratch = nullptr;  // this device's address of `scratch`
    hipStream_t stream = nullptr;
    size_t max_shared = 65536;
    // Monte Carlo launches: chain descriptors and per-chain counters
    DevBuf<McChain> mc_desc;
    DevBuf<unsigned long long> mc_cnt;
};

struct lg_system {
    lg_context* ctx = nullptr;
    size_t n = 0;
    // host mirrors of what the caller set
    std::vector<double> mass, charge;
    std::vector<uint32_t> kind;
    uint32_t nkinds = 1;
    double L[3] = {0, 0, 0};
    bool periodic = false;
    std::vector<std::pair<uint32_t, uint32_t>> bonds;
    std::map<std::pair<uint32_t, uint32_t>, PairSpec> pairs;
    std::map<std::pair<uint32_t, uint32_t>, BondedSpec> bond_pots;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, BondedSpec> angle_pots;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t>, BondedSpec> dihedral_pots;
    int coul_kind = COUL_NONE;
    double coul_rc = 0, coul_alpha = 0;
    int coul_kmax = 0, coul_restr = 0;
    double coul_s14 = 1.0;
    int precision = LG_PRECISION_FP64;
    lg_tuning tune{};  // 0 = automatic

    // derived topology
    std::vector<int> molid, mol_start;
    std::vector<std::tuple<int, int, int>> angles;
    std::vector<std::tuple<int, int, int, int>> dihedrals;
    bool topo_dirty = true, params_dirty = true, kspace_dirty = true, nlist_dirty = true, meta_dirty = true;

    // device state
    DevBuf<double4> pos, vel, acc, force, ref;
    DevBuf<int2> meta;
    // neighbour list
    Grid grid{};
    Box box{};
    double rlist = 0, skin = 0;
    int cap = 0;
    DevBuf<int> cellid, slot, count, start, tmp, sorted, nbr, nnbr, flags;  // flags[0] rebuild, [1] overflow
    DevBuf<float4> cpos;  // list-build candidate records in sorted order
    DevBuf<float4> frac;  // positions as the mixed pair kernel reads them (frac_pos)
    double3 frac_origin = make_double3(0.0, 0.0, 0.0);  // infinite cells: atom 0 at the last set
    // parameters
    DevBuf<PairParam> table;
    DevBuf<double> ptabs;  // TableComputation tables: energies then forces per kind pair
    bool gen_pairs = false;  // a pair potential other than LJ/Harmonic/Null, or a table
    DevBuf<int> pair_restr;
    DevBuf<double> pair_s14;
    std::vector<PairParam> h_table;
    CoulParam coul{};
    // molecules
    int nmol = 0, nterms = 0, max_mol = 1;
    DevBuf<int> d_mol_start;     // nmol + 1
    DevBuf<Term> terms;          // intra-molecular terms, sorted by type
    DevBuf<int> slot_start;      // per-atom CSR over contribution slots (n + 1)
    DevBuf<double4> contrib;     // one force contribution per (term, atom)
    DevBuf<BondedParam> bparam, aparam, dparam;
    // k-space
    int nk = 0, nrows = 0, nunits = 0, ksplit = 1;
    int tpa = 8;  // lanes per atom in k_pairs
    DevBuf<double4> fk;  // k-space partial forces [ksplit][n]
    DevBuf<int4> kidx;
    DevBuf<KRow> rows;
    DevBuf<KUnit> units;         // rho pass work units (up to 4 k-vectors of a row)
    DevBuf<int> kslot;           // k -> slot of the rho partials (j * nunits + unit)
    // matrix-core rho: column tiles of (iky, ikz), output -> k map, identity slots
    int mfma_nt = 0;
    DevBuf<int> mcols, mkmap, kslot_id;
    // erfcx tables of the real-space erfc (empty: library erfc)
    DevBuf<double> ecx64;
    DevBuf<float> ecx32;
    double ecx_inv_h = 0.0;
    // the mixed pair kernel's real-space force factor
    // h(r) = erfc(alpha r) + (2 alpha / sqrt(pi)) r exp(-alpha^2 r^2) on
    // [0, rc], FP32 pieces (hforce_ld); hf_ok when the fit met 1e-7
    DevBuf<float> hf32;
    double hf_inv_h = 0.0;
    bool hf_ok = false;
    bool ecx_ok = false;
    double kg[3] = {0, 0, 0};    // 2 pi / L per axis
    DevBuf<double> efac, vfac;
    DevBuf<KRec> krec;           // rho(k) and force factor per k
    DevBuf<double2> rho, rho_partial;
    DevBuf<double2> phase_tab;  // e^{i n k_axis x}, phase_rows(kmax) x ld
    DevBuf<double> phase_q;     // charges padded to ld
    DevBuf<float2> phase_tab32;  // mixed form: FP32 phase factors
    DevBuf<float> phase_q32;
    DevBuf<KRec32> krec32;       // mixed form: FP32 copy of krec
    int ld = 0;
    // energy partials
    DevBuf<double> epart_pairs, epart_mol, epart_k, ke_part, lambda;
    // when set, the energy partials go here instead (lg_potential_energy_noalloc)
    double* ep_pairs_at = nullptr;
    double* ep_mol_at = nullptr;
    double* ep_k_at = nullptr;
    // molecular virial: centres of mass and per-block partial sums
    DevBuf<double4> mol_com;
    DevBuf<double> molvir_part;
    // MD
    int integrator = LG_INTEGRATOR_VELOCITY_VERLET;
    double dt = 1.0, berendsen_T = 0.0, berendsen_tau = 1.0;
    bool thermostat = false;  // BerendsenThermostat(berendsen_T, berendsen_tau) after each step
    bool md_ready = false;
    // Monte Carlo energy cache (cache.rs:22-48): cached terms, rho(k) and
    // the staged move; cell list of the device chain
    bool mc_ready = false, mc_staged = false, mc_cells_valid = false;
    double c_pairs = 0, c_tail = 0, c_bonds = 0, c_angles = 0, c_coul = 0;
    uint32_t mc_mol = 0;
    double mc_dp = 0, mc_dc = 0;
    DevBuf<double2> mc_rho, mc_drho;
    DevBuf<double4> mc_newpos;
    DevBuf<double> mc_out;
    int mc_nc[3] = {0, 0, 0};
    double mc_csize[3] = {0, 0, 0};
    int mc_ncell = 0, mc_ccap = 0;
    DevBuf<double4> mc_cell_pos;
    DevBuf<int2> mc_cell_ids;
    DevBuf<int> mc_cell_count, mc_atom_cell;
    // snapshot for neighbour-list overflow recovery
    DevBuf<double4> snap_pos, snap_vel, snap_acc;
    // Verlet build deferred by launch_nlist into the rho(k) launch (k_cells_rho)
    bool cells_pending = false, cells_nearest = false, direct_pending = false;
    CellsArgs cells_args{};
    DirectArgs direct_args{};
    // MD chunk captured as a HIP graph, with what it was captured for
    hipGraphExec_t md_exec = nullptr;
    uint64_t md_key[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint64_t cfg_epoch = 0;  // bumped by prepare() whenever anything was re-derived
    // kernel timing (lg_profile_*)
    uint32_t prof_mask = 0;
    std::vector<hipEvent_t> prof_events;
    std::vector<std::pair<int, size_t>> prof_records;  // (class, first event index)
    size_t prof_used = 0;
};

namespace {

hipStream_t S(lg_system* s) { return s->ctx->stream; }

// Event bracket around a kernel class; returns the record index or -1.
long prof_start(lg_system* s, int cls) {
    if (!(s->prof_mask & (1u << cls))) return -1;
    while (s->prof_events.size() < s->prof_used + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        s->prof_events.push_back(e);
    }
    size_t idx = s->prof_used;
    s->prof_used += 2;
    if (hipEventRecord(s->prof_events[idx], s->ctx->stream) != hipSuccess) return -1;
    s->prof_records.push_back({cls, idx});
    return long(s->prof_records.size() - 1);
}

void prof_stop(lg_system* s, long rec) {
    if (rec < 0) return;
    (void)hipEventRecord(s->prof_events[s->prof_records[size_t(rec)].second + 1], s->ctx->stream);
}

// An ordinary launch (hipLaunchKernel) of `kernel` with its arguments
// converted to the kernel's own parameter types.
template <typename... KArgs, size_t... I>
void launch_plain(const void* kernel, dim3 grid, dim3 block, uint32_t shmem, hipStream_t st,
                  std::tuple<KArgs...>& a, std::index_sequence<I...>) {
    void* argv[sizeof...(KArgs) > 0 ? sizeof...(KArgs) : 1] = {static_cast<void*>(&std::get<I>(a))...};
    (void)hipLaunchKernel(kernel, grid, block, argv, shmem, st);
}

// A launch timed, when its class is profiled, by the kernel's own start and
// end timestamps (the events of hipExtLaunchKernelGGL), which agree with a
// kernel trace; an event pair recorded around the launch would add the
// dispatch of the event packets to every kernel's time.  Unprofiled, an
// ordinary launch: hipExtLaunchKernelGGL costs more host time per launch,
// which small systems (a step every ~30 us, five launches) notice.
template <typename... KArgs, typename... Args>
void launch_timed(lg_system* s, int cls, void (*kernel)(KArgs...), dim3 grid, dim3 block, uint32_t shmem,
                  Args... args) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!(s->prof_mask & (1u << cls))) {
        std::tuple<KArgs...> a(args...);
        launch_plain(reinterpret_cast<const void*>(kernel), grid, block, shmem, S(s), a,
                     std::index_sequence_for<KArgs...>{});
        return;
    }
    {
        bool ok = true;
        while (ok && s->prof_events.size() < s->prof_used + 2) {
            hipEvent_t e;
            ok = hipEventCreate(&e) == hipSuccess;
            if (ok) s->prof_events.push_back(e);
        }
        if (ok) {
            const size_t idx = s->prof_used;
            s->prof_used += 2;
            s->prof_records.push_back({cls, idx});
            e0 = s->prof_events[idx];
            e1 = s->prof_events[idx + 1];
        }
    }
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, S(s), e0, e1, 0u, args...);
}

int blocks_for(size_t n, int b) { return int((n + b - 1) / b); }

// Blocks of the pair part of a pair launch: rounded up to a multiple of 8,
// so that the kernels' XCD-contiguous block order (xcd_contiguous) is a
// permutation of them.
int pair_blocks_of(size_t n, int tpa) { return (blocks_for(n * size_t(tpa), kPairBlock) + 7) / 8 * 8; }

// ---- topology (bonding.rs:78-155, configuration.rs:124-144) ---------------------------
int build_topology(lg_system* s, std::vector<std::tuple<int, int, int>>& ipairs_out) {
    size_t n = s->n;
    std::vector<size_t> parent(n);
    std::iota(parent.begin(), parent.end(), size_t(0));
    std::function<size_t(size_t)> find = [&](size_t x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    for (auto& b : s->bonds) {
        size_t a = find(b.first), c = find(b.second);
        if (a != c) parent[std::max(a, c)] = std::min(a, c);
    }
    s->molid.assign(n, 0);
    s->mol_start.clear();
    int m = 0;
    for (size_t i = 0; i < n; i++) {
        size_t root = find(i);
        if (root == i) {
            s->mol_start.push_back(int(i));
            s->molid[i] = m++;
        } else {
            if (s->molid[root] != m - 1 || find(i - 1) != root)
                return fail(LG_E_INVALID, "molecules must be contiguous ranges of atoms");
            s->molid[i] = s->molid[root];
        }
    }
    s->mol_start.push_back(int(n));
    s->nmol = m;

    std::set<std::tuple<int, int, int>> angles;
    std::set<std::tuple<int, int, int, int>> dihedrals;
    auto make_angle = [](int a, int b, int c) { return std::make_tuple(std::min(a, c), b, std::max(a, c)); };
    auto make_dihedral = [](int a, int b, int c, int d) {
        if (std::max(a, b) < std::max(c, d)) return std::make_tuple(a, b, c, d);
        return std::make_tuple(d, c, b, a);
    };
    std::vector<std::pair<int, int>> bonds;
    for (auto& b : s->bonds) bonds.push_back({int(b.first), int(b.second)});
    // The angles are the pairs of bonds that share an atom, and the
    // dihedrals an angle and a third bond at one of its ends: only bonds at
    // an atom of the current ones can match, so each search runs over the
    // bonds incident to those atoms (a pass over all bonds for each bond
    // took ~2 s for 36 000-atom water).  The sets make the result
    // independent of the visiting order.
    std::vector<std::vector<int>> inc(n);
    for (size_t k = 0; k < bonds.size(); k++) {
        inc[size_t(bonds[k].first)].push_back(int(k));
        if (bonds[k].second != bonds[k].first) inc[size_t(bonds[k].second)].push_back(int(k));
    }
    // each bond at atom x or atom y, once
    auto incident = [&](int x, int y, auto&& fn) {
        for (int k : inc[size_t(x)]) fn(bonds[size_t(k)]);
        if (y != x)
            for (int k : inc[size_t(y)])
                if (bonds[size_t(k)].first != x && bonds[size_t(k)].second != x) fn(bonds[size_t(k)]);
    };
    for (auto& b1 : bonds)
        incident(b1.first, b1.second, [&](const std::pair<int, int>& b2) {
            if (b1 == b2) return;
            std::tuple<int, int, int> angle;
            if (b1.first == b2.second)
                angle = make_angle(b2.first, b2.second, b1.second);
            else if (b1.second == b2.first)
                angle = make_angle(b1.first, b1.second, b2.second);
            else if (b1.second == b2.second)
                angle = make_angle(b1.first, b1.second, b2.first);
            else if (b1.first == b2.first)
                angle = make_angle(b1.second, b1.first, b2.second);
            else
                return;
            angles.insert(angle);
            const int ai = std::get<0>(angle), aj = std::get<1>(angle), ak = std::get<2>(angle);
            incident(ak, ai, [&](const std::pair<int, int>& b3) {
                if (b2 == b3) return;
                if (ak == b3.first && aj != b3.second)
                    dihedrals.insert(make_dihedral(ai, aj, ak, b3.second));
                else if (ak == b3.second && aj != b3.first)
                    dihedrals.insert(make_dihedral(ai, aj, ak, b3.first));
                else if (ai == b3.second && aj != b3.first)
                    dihedrals.insert(make_dihedral(b3.first, ai, aj, ak));
                else if (ai == b3.first && aj != b3.second)
                    dihedrals.insert(make_dihedral(b3.second, ai, aj, ak));
            });
        });
    s->angles.assign(angles.begin(), angles.end());
    s->dihedrals.assign(dihedrals.begin(), dihedrals.end());
    std::map<std::pair<int, int>, int> flags;
    for (auto& [i, j] : bonds) {
        flags[{i, j}] |= 1;
        flags[{j, i}] |= 1;
    }
    for (auto& [i, j, k] : s->angles) {
        flags[{i, k}] |= 2;
        flags[{k, i}] |= 2;
    }
    for (auto& [i, j, k, mm] : s->dihedrals) {
        flags[{i, mm}] |= 4;
        flags[{mm, i}] |= 4;
    }
    ipairs_out.clear();
    for (int mol = 0; mol < s->nmol; mol++) {
        int m0 = s->mol_start[mol], m1 = s->mol_start[mol + 1];
        for (int i = m0; i < m1; i++)
            for (int j = i + 1; j < m1; j++) {
                auto it = flags.find({i, j});
                int f = it == flags.end() ? 0 : it->second;
                int path = (f & 1) ? BP_ONE : (f & 2) ? BP_TWO : (f & 4) ? BP_THREE : BP_FAR;
                ipairs_out.push_back({i, j, path});
            }
    }
    return LG_OK;
}

template <class T>
int upload(DevBuf<T>& b, const std::vector<T>& v, hipStream_t st) {
    LG_TRY(b.ensure(v.size()));
    if (!v.empty()) LG_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return LG_OK;
}

// Intra-molecular terms in type order (same-molecule pairs, bonds, angles,
// dihedrals), each with one contribution slot per atom.  An atom's slots are
// numbered in term order, which fixes the order in which its contributions
// are added.
int prepare_topology(lg_system* s) {
    std::vector<std::tuple<int, int, int>> ip;
    LG_TRY(build_topology(s, ip));
    hipStream_t st = S(s);
    std::vector<BondedParam> h_bp, h_ap, h_dp;
    auto kind_of = [&](int a) { return s->kind[a]; };
    auto bonded = [](const BondedSpec& b) {
        return BondedParam{b.pot, 0, 0, b.p[0], b.p[1], b.p[2], b.p[3], b.p[4], 0.0};
    };
    const BondedParam none{POT_NONE, 0, 0, 0, 0, 0, 0, 0, 0};
    auto find_bond = [&](int a, int b) -> BondedParam {
        uint32_t x = kind_of(a), y = kind_of(b);
        auto it = s->bond_pots.find({std::min(x, y), std::max(x, y)});
        if (it == s->bond_pots.end()) return none;
        return bonded(it->second);
    };
    auto find_angle = [&](int a, int b, int c) -> BondedParam {
        uint32_t x = kind_of(a), y = kind_of(b), z = kind_of(c);
        if (x > z) std::swap(x, z);
        auto it = s->angle_pots.find({x, y, z});
        if (it == s->angle_pots.end()) return none;
        return bonded(it->second);
    };
    auto dkey = [](uint32_t i, uint32_t j, uint32_t k, uint32_t m) {
        uint32_t ij = std::max(i, j), km = std::max(k, m);
        if (ij == km) {
            if (std::min(i, j) < std::min(k, m)) return std::make_tuple(i, j, k, m);
            return std::make_tuple(m, k, j, i);
        }
        if (ij < km) return std::make_tuple(i, j, k, m);
        return std::make_tuple(m, k, j, i);
    };
    auto find_dihedral = [&](int a, int b, int c, int d) -> BondedParam {
        auto it = s->dihedral_pots.find(dkey(kind_of(a), kind_of(b), kind_of(c), kind_of(d)));
        if (it == s->dihedral_pots.end()) return none;
        return bonded(it->second);
    };
    std::vector<Term> h_terms;
    h_terms.reserve(ip.size() + s->bonds.size() + s->angles.size() + s->dihedrals.size());
    auto term = [](int type, int param, int a0, int a1, int a2, int a3) {
        Term t;
        t.atoms = make_int4(a0, a1, a2, a3);
        t.slots = make_int4(-1, -1, -1, -1);
        t.type = type;
        t.param = param;
        return t;
    };
    for (auto& [i, j, path] : ip) h_terms.push_back(term(TERM_PAIR, path, i, j, -1, -1));
    for (auto& b : s->bonds) {
        int i = int(b.first), j = int(b.second);
        h_bp.push_back(find_bond(i, j));
        h_terms.push_back(term(TERM_BOND, int(h_bp.size() - 1), i, j, -1, -1));
    }
    for (auto& [i, j, k] : s->angles) {
        h_ap.push_back(find_angle(i, j, k));
        h_terms.push_back(term(TERM_ANGLE, int(h_ap.size() - 1), i, j, k, -1));
    }
    for (auto& [i, j, k, m] : s->dihedrals) {
        h_dp.push_back(find_dihedral(i, j, k, m));
        h_terms.push_back(term(TERM_DIHEDRAL, int(h_dp.size() - 1), i, j, k, m));
    }
    // slots: count per atom, prefix sum, then hand out in term order
    std::vector<int> slot_start(s->n + 1, 0);
    auto natoms_of = [](const Term& t) { return t.type == TERM_DIHEDRAL ? 4 : t.type == TERM_ANGLE ? 3 : 2; };
    for (auto& t : h_terms) {
        const int* at = &t.atoms.x;
        for (int q = 0; q < natoms_of(t); q++) slot_start[at[q] + 1]++;
    }
    for (size_t i = 0; i < s->n; i++) slot_start[i + 1] += slot_start[i];
    std::vector<int> next(slot_start.begin(), slot_start.end() - 1);
    for (auto& t : h_terms) {
        const int* at = &t.atoms.x;
        int* sl = &t.slots.x;
        for (int q = 0; q < natoms_of(t); q++) sl[q] = next[at[q]]++;
    }
    s->nterms = int(h_terms.size());
    s->max_mol = 1;
    for (int m = 0; m < s->nmol; m++) s->max_mol = std::max(s->max_mol, s->mol_start[m + 1] - s->mol_start[m]);
    LG_TRY(upload(s->d_mol_start, s->mol_start, st));
    LG_TRY(upload(s->terms, h_terms, st));
    LG_TRY(upload(s->slot_start, slot_start, st));
    LG_TRY(s->contrib.ensure(std::max<size_t>(1, size_t(slot_start[s->n]))));
    LG_TRY(upload(s->bparam, h_bp, st));
    LG_TRY(upload(s->aparam, h_ap, st));
    LG_TRY(upload(s->dparam, h_dp, st));
    LG_TRY(s->epart_mol.ensure(size_t(std::max(1, blocks_for(s->nterms, 256))) * 11));
    s->meta_dirty = true;
    s->topo_dirty = false;
    return LG_OK;
}

int prepare_meta(lg_system* s) {
    std::vector<int2> m(s->n);
    for (size_t i = 0; i < s->n; i++) m[i] = make_int2(int(s->kind[i]), s->molid.empty() ? int(i) : s->molid[i]);
    LG_TRY(upload(s->meta, m, S(s)));
    s->meta_dirty = false;
    return LG_OK;
}

// A function f on [0, xmax] as kErfcxPieces polynomials of degree `deg` in
// t in [-1, 1] per piece: the piece's Chebyshev interpolant (deg + 1
// Chebyshev nodes, long double), expanded in powers of t.  Returns the
// largest relative error against f (long double) over 64 points per piece.
// Used for erfcx(x) = exp(x^2) erfc(x) and for the real-space force factor
// of the mixed pair kernel (hforce_ld).
static long double erfcx_ld(long double x) { return erfcl(x) * expl(x * x); }
template <typename T, typename F>
static double fit_pieces(double xmax, int deg, F fn, std::vector<T>& out) {
    const int M = kErfcxPieces, D = deg;
    const long double pi = 3.141592653589793238462643383279502884L;
    const long double h = (long double)xmax / M;
    out.assign(size_t(M) * (D + 1), T(0));
    double worst = 0.0;
    std::vector<long double> f(D + 1), c(D + 1), mono(D + 1), tprev(D + 1), tcur(D + 1), tnext(D + 1);
    for (int k = 0; k < M; k++) {
        const long double mid = (k + 0.5L) * h, half = 0.5L * h;
        for (int j = 0; j <= D; j++) f[j] = fn(mid + half * cosl(pi * (j + 0.5L) / (D + 1)));
        for (int m = 0; m <= D; m++) {
            long double acc = 0;
            for (int j = 0; j <= D; j++) acc += f[j] * cosl(m * pi * (j + 0.5L) / (D + 1));
            c[m] = acc * 2.0L / (D + 1);
        }
        c[0] *= 0.5L;
        // sum c_m T_m(t) in powers of t (T_0 = 1, T_1 = t, T_{m+1} = 2t T_m - T_{m-1})
        std::fill(mono.begin(), mono.end(), 0.0L);
        std::fill(tprev.begin(), tprev.end(), 0.0L);
        std::fill(tcur.begin(), tcur.end(), 0.0L);
        tprev[0] = 1.0L;
        mono[0] += c[0];
        if (D >= 1) {
            tcur[1] = 1.0L;
            mono[1] += c[1];
        }
        for (int m = 1; m < D; m++) {
            std::fill(tnext.begin(), tnext.end(), 0.0L);
            for (int q = 0; q < D; q++) tnext[q + 1] += 2.0L * tcur[q];
            for (int q = 0; q <= D; q++) tnext[q] -= tprev[q];
            for (int q = 0; q <= D; q++) mono[q] += c[m + 1] * tnext[q];
            tprev = tcur;
            tcur = tnext;
        }
        for (int q = 0; q <= D; q++) out[size_t(k) * (D + 1) + q] = T(mono[q]);
        for (int z = 0; z < 64; z++) {
            const long double t = -1.0L + 2.0L * (z + 0.5L) / 64.0L;
            long double p = (long double)out[size_t(k) * (D + 1) + D];
            for (int q = D - 1; q >= 0; q--) p = p * t + (long double)out[size_t(k) * (D + 1) + q];
            const long double ref = fn(mid + half * t);
            worst = std::max(worst, double(fabsl(p - ref) / ref));
        }
    }
    return worst;
}
// Fitted tables by (function, its parameter, range, degree), kept for the
// process: systems with the same Ewald/Wolf alpha and cutoff (a serving
// stream of one water model) fit once.  A fit is a pure function of the key,
// so a cached table is the table the fit would return.
template <typename T, typename F>
static double fit_cached(int fn_id, double param, double xmax, int deg, F fn, std::vector<T>& out) {
    using Key = std::tuple<int, uint64_t, uint64_t, int>;
    static std::mutex mu;
    static std::map<Key, std::pair<std::vector<T>, double>> cache;
    uint64_t pb, xb;
    std::memcpy(&pb, &param, sizeof(pb));
    std::memcpy(&xb, &xmax, sizeof(xb));
    const Key key{fn_id, pb, xb, deg};
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            out = it->second.first;
            return it->second.second;
        }
    }
    const double err = fit_pieces(xmax, deg, fn, out);
    std::lock_guard<std::mutex> lk(mu);
    cache.emplace(key, std::make_pair(out, err));
    return err;
}
template <typename T>
static double fit_erfcx(double xmax, int deg, std::vector<T>& out) {
    return fit_cached(0, 0.0, xmax, deg, erfcx_ld, out);
}


int prepare_params(lg_system* s) {
    uint32_t K = s->nkinds;
    if (K > uint32_t(kMaxKinds)) return fail(LG_E_UNSUPPORTED, "more than 16 particle kinds");
    s->h_table.assign(size_t(K) * K, PairParam{POT_NONE, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0});
    std::vector<int> restr(size_t(K) * K, 0);
    std::vector<double> s14(size_t(K) * K, 1.0);
    std::vector<double> tabs;
    s->gen_pairs = false;
    for (auto& [key, p] : s->pairs) {
        int tab_off = 0;
        if (p.v.tab_n > 0) {
            std::vector<double> te, tf;
            host_table(p.v, te, tf);
            tab_off = int(tabs.size());
            tabs.insert(tabs.end(), te.begin(), te.end());
            tabs.insert(tabs.end(), tf.begin(), tf.end());
        }
        if (!pot_is_fast(p.v.pot, int(p.v.tab_n))) s->gen_pairs = true;
        for (int o = 0; o < 2; o++) {
            uint32_t a = o ? key.second : key.first, b = o ? key.first : key.second;
            if (a >= K || b >= K) return fail(LG_E_INVALID, "pair potential kind out of range");
            PairParam& pp = s->h_table[size_t(a) * K + b];
            pp.pot = p.v.pot;
            pp.inter_excl = p.restriction == LG_RESTRICT_INTRA ? 1 : 0;
            pp.p0 = p.v.p[0];
            pp.p1 = p.v.p[1];
            pp.p2 = p.v.p[2];
            pp.p3 = p.v.p[3];
            pp.p4 = p.v.p[4];
            pp.rc = p.rc;
            pp.tab_n = int(p.v.tab_n);
            pp.tab_off = tab_off;
            pp.tab_delta = p.v.tab_n > 0 ? p.v.tab_max / double(p.v.tab_n) : 0.0;
            // pairs.rs:86-95: shift = V(cutoff) of the potential (of its
            // table when tabulated)
            pp.shift = p.shifted ? host_energy(p.v, p.rc) : 0.0;
            restr[size_t(a) * K + b] = p.restriction;
            s14[size_t(a) * K + b] = p.scale14;
        }
    }
    if (tabs.empty()) tabs.push_back(0.0);
    LG_TRY(upload(s->ptabs, tabs, S(s)));
    LG_TRY(upload(s->table, s->h_table, S(s)));
    LG_TRY(upload(s->pair_restr, restr, S(s)));
    LG_TRY(upload(s->pair_s14, s14, S(s)));
    CoulParam c{};
    c.kind = s->coul_kind;
    c.inter_excl = s->coul_restr == LG_RESTRICT_INTRA ? 1 : 0;
    c.rc = s->coul_rc;
    if (s->coul_kind == COUL_EWALD) {
        c.alpha = s->coul_alpha;
    } else if (s->coul_kind == COUL_WOLF) {
        // wolf.rs:68-84
        c.alpha = kPi / s->coul_rc;
        double ac = c.alpha * s->coul_rc;
        double ac2 = ac * ac;
        c.e_const = std::erfc(ac) / s->coul_rc;
        c.f_const = std::erfc(ac) / (s->coul_rc * s->coul_rc) + kFrac2SqrtPi * c.alpha * std::exp(-ac2) / s->coul_rc;
    }
    s->coul = c;
    // erfcx tables for the pair kernels' real-space erfc (Ewald and Wolf);
    // kept only if the fit meets 1e-14 (FP64) and 1e-7 (FP32) relative
    s->ecx_ok = false;
    s->hf_ok = false;
    if ((s->coul_kind == COUL_EWALD || s->coul_kind == COUL_WOLF) && c.alpha > 0.0 && c.rc > 0.0) {
        const double xmax = c.alpha * c.rc * (1.0 + 1e-9);
        std::vector<double> t64;
        std::vector<float> t32;
        const double e64 = fit_erfcx(xmax, kErfcxDeg64, t64);
        const double e32 = fit_erfcx(xmax, kErfcxDeg32, t32);
        if (xmax < 27.0 && e64 < 1e-14 && e32 < 1e-7) {
            LG_TRY(upload(s->ecx64, t64, S(s)));
            LG_TRY(upload(s->ecx32, t32, S(s)));
            s->ecx_inv_h = kErfcxPieces / xmax;
            s->ecx_ok = true;
        }
        // h(r): erfc(alpha r) r^-3 + g0 exp(-alpha^2 r^2) r^-2 = h(r) r^-3,
        // the Ewald real-space force over r (and the r-dependent part of
        // Wolf's), in one table lookup without an exponential
        const long double al = c.alpha, g0 = 2.0L * al / sqrtl(3.141592653589793238462643383279502884L);
        auto hforce_ld = [al, g0](long double r) { return erfcl(al * r) + g0 * r * expl(-al * al * r * r); };
        const double rmax = c.rc * (1.0 + 1e-9);
        std::vector<float> th;
        const double eh = fit_cached(1, c.alpha, rmax, kErfcxDeg32, hforce_ld, th);
        s->hf_ok = false;
        if (s->ecx_ok && eh < 1e-7) {
            LG_TRY(upload(s->hf32, th, S(s)));
            s->hf_inv_h = kErfcxPieces / rmax;
            s->hf_ok = true;
        }
    }
    s->params_dirty = false;
    s->nlist_dirty = true;
    return LG_OK;
}

// ewald.rs:130-183 and :353-378 for an orthorhombic cell
// Launch shape of the matrix-core rho pass: atoms per stage and tiles per
// wave, the first candidate with at least 256 blocks (one per CU) -- wider
// stages and more tiles per wave share more staging and operand work --
// then stages merged per block so that the grid stays near 1024-2048 blocks
// and the partials (chunks x k-vectors) few.
struct RhoShape {
    int width, wlog, cpb, ngroups, nchunks;
};
// dynamic LDS of a matrix-core rho launch: its own, or the Verlet build's
// when the two share the launch and the buffer
static size_t rho_mfma_bytes(int width, int kmax, bool mixed) {
    const size_t r = mixed ? rho_mfma_lds<float>(width, kmax) : rho_mfma_lds<double>(width, kmax);
    return std::max(r, sizeof(CellsLds));
}
// Launch shape of the matrix-core rho pass: the widest stage (atoms staged
// per block: wider shares more staging per k) that still gives at least 256
// blocks (one per CU) and fits the LDS, then stages merged per block so that
// the grid stays near 1024-2048 blocks and the partials (chunks x k-vectors)
// few.
static RhoShape rho_mfma_shape(const lg_system* s, bool mixed) {
    // widths 16..128 (FP32) or 16..64 (FP64): the kernel's trips take four
    // atoms of each of its four sub-waves
    const int wl0 = mixed ? 7 : 6, wlmin = 4;
    RhoShape sh{};
    sh.ngroups = (s->mfma_nt + 2 * kRhoMfmaTpw - 1) / (2 * kRhoMfmaTpw);
    sh.wlog = wlmin;
    for (int wl = wl0; wl >= wlmin; wl--) {
        if (rho_mfma_bytes(1 << wl, s->coul_kmax, mixed) > s->ctx->max_shared) continue;
        sh.wlog = wl;
        if (size_t(blocks_for(s->ld, 1 << wl)) * sh.ngroups >= 256) break;
    }
    sh.width = 1 << sh.wlog;
    const int stages = blocks_for(s->ld, sh.width);
    sh.cpb = std::max(1, int(size_t(stages) * sh.ngroups / 1024));
    sh.nchunks = (stages + sh.cpb - 1) / sh.cpb;
    return sh;
}

int prepare_kspace(lg_system* s) {
    s->kspace_dirty = false;
    s->nk = 0;
    s->nrows = 0;
    if (s->coul_kind != COUL_EWALD) return LG_OK;
    if (!s->periodic) return fail(LG_E_INVALID, "Ewald is not defined with infinite unit cell");
    double det = s->L[0] * (s->L[1] * s->L[2] - 0.0 * 0.0) - 0.0 + 0.0;
    double inv_det = 1.0 / det;
    double inv[3] = {(s->L[1] * s->L[2] - 0.0 * 0.0) * inv_det, (s->L[0] * s->L[2] - 0.0 * 0.0) * inv_det,
                     (s->L[0] * s->L[1] - 0.0 * 0.0) * inv_det};
    auto kvec = [&](double a, double b, double c, double out[3]) {
        out[0] = 2.0 * kPi * inv[0] * a;
        out[1] = 2.0 * kPi * inv[1] * b;
        out[2] = 2.0 * kPi * inv[2] * c;
    };
    double k111[3];
    kvec(1, 1, 1, k111);
    int kmax = s->coul_kmax;
    double mx = std::max(k111[0], std::max(k111[1], k111[2])) * double(kmax);
    double kmax2 = 1.0001 * mx * mx;
    double alpha = s->coul_alpha;
    double a4 = 0.25 / (alpha * alpha);
    double volume = s->L[0] * s->L[1] * s->L[2];
    double four_pi_v = 4.0 * kPi / volume;
    std::vector<int4> idx;
    std::vector<double> efac;
    std::vector<KRec> rec;
    std::vector<double> vf[6];
    auto push = [&](int x, int y, int z) {
        double k[3];
        kvec(x, y, z, k);
        double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        if (k2 > kmax2) return;
        double energy = four_pi_v * std::exp(-k2 * a4) / k2;
        idx.push_back(make_int4(x, y, z, 0));
        efac.push_back(energy);
        rec.push_back(KRec{0.0, 0.0, 2.0 * energy, (2.0 * energy) * double(z)});
        double f = -2.0 * (1.0 / k2 + a4);
        // virial = energy * (I + f k (x) k); components xx yy zz xy xz yz
        vf[0].push_back(energy * (1.0 + f * (k[0] * k[0])));
        vf[1].push_back(energy * (1.0 + f * (k[1] * k[1])));
        vf[2].push_back(energy * (1.0 + f * (k[2] * k[2])));
        vf[3].push_back(energy * (0.0 + f * (k[0] * k[1])));
        vf[4].push_back(energy * (0.0 + f * (k[0] * k[2])));
        vf[5].push_back(energy * (0.0 + f * (k[1] * k[2])));
    };
    for (int x = 1; x < kmax; x++)
        for (int y = -kmax; y < kmax; y++)
            for (int z = -kmax; z < kmax; z++) push(x, y, z);
    for (int y = 1; y < kmax; y++)
        for (int z = -kmax; z < kmax; z++) push(0, y, z);
    for (int z = 1; z < kmax; z++) push(0, 0, z);
    s->nk = int(idx.size());
    std::vector<KRow> rows;
    for (int k = 0; k < s->nk; k++) {
        if (rows.empty() || rows.back().ikx != idx[k].x || rows.back().iky != idx[k].y ||
            idx[k - 1].z + 1 != idx[k].z)
            rows.push_back(KRow{short(idx[k].x), short(idx[k].y), short(idx[k].z), 0, k, 0});
        rows.back().count++;
    }
    s->nrows = int(rows.size());
    std::vector<KUnit> units;
    for (const KRow& r : rows)
        for (int j = 0; j < r.count; j += kRhoUnitK)
            units.push_back(
                KUnit{r.ikx, r.iky, short(r.ikz0 + j), short(std::min(kRhoUnitK, r.count - j)), r.kbase + j, 0});
    s->nunits = int(units.size());
    std::vector<int> kslot(s->nk);
    for (int u = 0; u < s->nunits; u++)
        for (int j = 0; j < units[u].nv; j++) kslot[units[u].kb + j] = j * s->nunits + u;
    for (int d = 0; d < 3; d++) s->kg[d] = 2.0 * kPi * inv[d];
    // matrix-core rho: distinct (iky, ikz) columns in tiles of 16; rows ikx
    // 0..15 cover the set when kmax <= 16
    std::vector<int> mcols, mkmap, kid(s->nk);
    s->mfma_nt = 0;
    if (kmax <= 16) {
        std::vector<std::pair<int, int>> cl;
        for (int k = 0; k < s->nk; k++) cl.push_back({idx[k].y, idx[k].z});
        std::sort(cl.begin(), cl.end());
        cl.erase(std::unique(cl.begin(), cl.end()), cl.end());
        const int nt = int((cl.size() + 15) / 16);
        mcols.assign(size_t(nt) * 16, (kmax << 16) | kmax);
        for (size_t j = 0; j < cl.size(); j++) mcols[j] = ((cl[j].first + kmax) << 16) | (cl[j].second + kmax);
        mkmap.assign(size_t(nt) * 256, -1);
        for (int k = 0; k < s->nk; k++) {
            int j = int(std::lower_bound(cl.begin(), cl.end(), std::make_pair(idx[k].y, idx[k].z)) - cl.begin());
            mkmap[size_t(j / 16) * 256 + idx[k].x * 16 + j % 16] = k;
        }
        s->mfma_nt = nt;
    }
    for (int k = 0; k < s->nk; k++) kid[k] = k;
    std::vector<double> vflat;
    for (int q = 0; q < 6; q++) vflat.insert(vflat.end(), vf[q].begin(), vf[q].end());
    hipStream_t st = S(s);
    LG_TRY(upload(s->kidx, idx, st));
    LG_TRY(upload(s->rows, rows, st));
    LG_TRY(upload(s->units, units, st));
    LG_TRY(upload(s->kslot, kslot, st));
    if (s->mfma_nt > 0) {
        LG_TRY(upload(s->mcols, mcols, st));
        LG_TRY(upload(s->mkmap, mkmap, st));
        LG_TRY(upload(s->kslot_id, kid, st));
    }
    LG_TRY(upload(s->efac, efac, st));
    LG_TRY(upload(s->vfac, vflat, st));
    LG_TRY(upload(s->krec, rec, st));
    std::vector<KRec32> r32(rec.size());
    for (size_t k = 0; k < rec.size(); k++)
        r32[k] = KRec32{0.f, 0.f, float(rec[k].g), float(rec[k].gz)};
    LG_TRY(upload(s->krec32, r32, st));
    LG_TRY(s->rho.ensure(s->nk));
    if (kmax > 1000) return fail(LG_E_UNSUPPORTED, "kmax too large");
    // enough waves to fill the device (~4096: 64 atoms x 4 waves per block,
    // grid.y row splits), with as few force partials as that allows
    int fwaves = 4 * blocks_for(s->n, 64);
    s->ksplit = std::max(1, std::min(32, std::min(s->nrows, (4096 + fwaves - 1) / fwaves)));
    if (s->tune.kspace_split > 0) s->ksplit = std::min(s->tune.kspace_split, std::max(1, s->nrows));
    LG_TRY(s->fk.ensure(size_t(s->ksplit) * s->n));
    s->ld = blocks_for(s->n, 64) * 64;
    // rho partials: the row form's narrowest atom chunk (64) or the
    // matrix-core form's chunks, whichever needs more
    size_t npart = size_t(blocks_for(s->ld, 64)) * kRhoUnitK * s->nunits;
    if (s->mfma_nt > 0)
        for (int m = 0; m < 2; m++) npart = std::max(npart, size_t(rho_mfma_shape(s, m == 1).nchunks) * s->nk);
    LG_TRY(s->rho_partial.ensure(npart));
    // the staged row kernel loads kRhoStageRows rows per round (the rows
    // past the phase rows are loaded and not used)
    const size_t trows = size_t(std::max(phase_rows(kmax), kRhoStageRows));
    LG_TRY(s->phase_tab.ensure(trows * s->ld));
    LG_TRY(s->phase_q.ensure(s->ld));
    LG_TRY(s->phase_tab32.ensure(trows * s->ld));
    LG_TRY(s->phase_q32.ensure(s->ld));
    LG_TRY(s->epart_k.ensure(size_t(blocks_for(s->nk, kRhoReduceK)) * 7));
    return LG_OK;
}

double max_cutoff(lg_system* s) {
    double rc = 0.0;
    for (auto& [k, p] : s->pairs)
        if (p.v.pot != LG_POT_NULL) rc = std::max(rc, p.rc);
    if (s->coul_kind != COUL_NONE) rc = std::max(rc, s->coul_rc);
    return rc;
}

int prepare_nlist(lg_system* s) {
    size_t n = s->n;
    double rc = max_cutoff(s);
    s->skin = rc > 0 ? 0.08 * rc : 0.0;
    s->rlist = rc + s->skin;
    Box& b = s->box;
    b.periodic = s->periodic ? 1 : 0;
    for (int d = 0; d < 3; d++) {
        b.L[d] = s->periodic ? s->L[d] : 0.0;
        b.inv[d] = s->periodic ? 1.0 / s->L[d] : 0.0;
    }
    // cells of edge >= rlist / 2 searched 2 deep when 5 fit on an axis, else
    // >= rlist searched 1 deep when 3 fit, else one cell (nearest image); a
    // sparse box whose half cells would outnumber 8 per atom (and 2^22) uses
    // full cells, at most 256 per axis
    Grid& g = s->grid;
    auto layout = [&](bool half) {
        for (int d = 0; d < 3; d++) {
            int nc = 1, reach = 0;
            if (s->periodic && s->rlist > 0) {
                int nh = int(std::floor(s->L[d] / (0.5 * s->rlist * (1.0 + 1e-12))));
                int nf = int(std::floor(s->L[d] / (s->rlist * (1.0 + 1e-12))));
                if (half && nh >= 5) {
                    nc = nh;
                    reach = 2;
                } else if (nf >= 3) {
                    nc = std::min(nf, 256);
                    reach = 1;
                }
            }
            g.n[d] = nc;
            g.reach[d] = reach;
            g.size[d] = s->periodic ? s->L[d] / nc : 0.0;
            g.inv_size[d] = s->periodic ? 1.0 / g.size[d] : 0.0;
        }
    };
    layout(true);
    double cells = double(g.n[0]) * g.n[1] * g.n[2];
    if (cells > std::min(8.0 * double(n) + 4096.0, double(1 << 22))) layout(false);
    g.ncell = g.n[0] * g.n[1] * g.n[2];
    // capacity from the mean density, with margin; exact sizes are checked
    // after every build (overflow flag) and grown on demand.
    size_t capacity = n;
    if (s->periodic && s->rlist > 0) {
        double volume = s->L[0] * s->L[1] * s->L[2];
        double expect = 4.0 / 3.0 * kPi * s->rlist * s->rlist * s->rlist * double(n) / volume;
        capacity = size_t(std::ceil(expect * 1.5 + 64));
        capacity = std::min(capacity, n);
    }
    if (s->cap > int(capacity)) capacity = size_t(s->cap);
    s->cap = int(std::max<size_t>(capacity, 1));
    LG_TRY(s->cellid.ensure(n));
    LG_TRY(s->slot.ensure(n));
    LG_TRY(s->count.ensure(g.ncell));
    LG_TRY(s->start.ensure(g.ncell + 1));
    LG_TRY(s->tmp.ensure(n));
    LG_TRY(s->sorted.ensure(n));
    LG_TRY(s->cpos.ensure(n));
    LG_TRY(s->frac.ensure(n));
    LG_TRY(s->nbr.ensure(size_t(s->cap) * n));
    LG_TRY(s->nnbr.ensure(n));
    LG_TRY(s->flags.ensure(4));
    LG_TRY(s->ref.ensure(n));
    // lanes per atom: enough waves to fill 256 CUs, fewer lanes for long
    // lists; not fewer than 8 (1M-atom Kob-Andersen: 4 lanes 660/720 us per
    // mixed/FP64 step, 8 lanes 562/625 us; lumol_amd.tuning picks per bucket)
    s->tpa = n <= 8192 ? 64 : n <= 32768 ? 32 : n <= 131072 ? 16 : 8;
    if (s->tune.pair_lanes > 0) s->tpa = s->tune.pair_lanes;
    if (s->gen_pairs) s->tpa = 16;  // the general pair kernel has one instance
    LG_TRY(s->epart_pairs.ensure(size_t(pair_blocks_of(n, s->tpa)) * 8));
    int init[4] = {1, 0, 0, 0};  // force a rebuild; clear overflow
    LG_HIP(hipMemcpyAsync(s->flags.p, init, sizeof(init), hipMemcpyHostToDevice, S(s)));
    LG_HIP(hipMemsetAsync(s->count.p, 0, g.ncell * sizeof(int), S(s)));
    s->nlist_dirty = false;
    return LG_OK;
}


int prepare(lg_system* s) {
    if (s->n == 0) return fail(LG_E_INVALID, "empty system");
    if (s->topo_dirty || s->meta_dirty || s->params_dirty || s->kspace_dirty || s->nlist_dirty) {
        s->cfg_epoch++;
        s->mc_ready = false;
        s->mc_staged = false;
        s->mc_cells_valid = false;
    }
    if (s->topo_dirty) LG_TRY(prepare_topology(s));
    if (s->meta_dirty) LG_TRY(prepare_meta(s));
    if (s->params_dirty) LG_TRY(prepare_params(s));
    if (s->kspace_dirty) LG_TRY(prepare_kspace(s));
    if (s->periodic) {
        double half = std::min(s->L[0], std::min(s->L[1], s->L[2])) / 2.0;
        for (auto& [k, p] : s->pairs)
            if (p.rc > half)
                return fail(LG_E_INVALID, "The cutoff for this pair potential is bigger than half the cell length");
        // the cell may have changed after the coulomb potential was set
        if (s->coul_kind != COUL_NONE && s->coul_rc > half)
            return fail(LG_E_INVALID, "The cutoff for the coulomb potential is bigger than half the cell length");
    }
    if (s->nlist_dirty) LG_TRY(prepare_nlist(s));
    return LG_OK;
}

// ---- kernel sequences --------------------------------------------------------------------
// Neighbour-list maintenance.  `check` runs the displacement test here; the
// MD step does it inside the integrator's first half instead.  Every kernel
// returns at once when flags[0] (rebuild) is clear; the build kernel clears
// it again.
// Which list-build kernel runs: one workgroup per cell (k_nlist_cells) when
// there are enough cells to fill the device and enough atoms per cell to keep
// its waves busy, else TPA lanes per atom (k_nlist_build); the tuning knob
// nlist_build overrides (1 lanes per atom, 2 cells).
// Verlet build: 1 lanes per atom over the cell list, 2 one workgroup per
// cell over the cell list, 3 workgroups per cell (direct_split) over the
// atoms' cell tags (no cell list; up to 32768 atoms and 1023 cells per axis).  Default:
// the per-cell builds when there are >= 128 cells and >= 8 atoms per cell,
// the direct one where it applies.
static bool direct_fits(const lg_system* s) {
    return s->n <= 32768 && s->grid.n[0] <= 1023 && s->grid.n[1] <= 1023 && s->grid.n[2] <= 1023;
}
static bool cells_dense(const lg_system* s) {
    return s->grid.ncell >= 128 && s->n >= size_t(8) * s->grid.ncell;
}
static bool direct_build(const lg_system* s) {
    if (!direct_fits(s)) return false;
    if (s->tune.nlist_build == 3) return true;
    return s->tune.nlist_build == 0 && cells_dense(s);
}
static bool cell_build(const lg_system* s) {
    if (s->tune.nlist_build == 1) return false;
    if (s->tune.nlist_build == 2 || s->tune.nlist_build == 3) return true;
    return cells_dense(s);
}
// Where the displacement tests write the atoms' cell tags: only for the
// direct build.
static CellTag cell_tags(lg_system* s) {
    return CellTag{direct_build(s) ? s->cellid.p : nullptr, s->grid, s->box,
                   s->precision == LG_PRECISION_MIXED ? s->frac.p : nullptr, s->frac_origin};
}
// Workgroups per cell of the direct build: about 256 workgroups in all (one
// per CU of the MI355X), at most 8 per cell.
static int direct_split(const lg_system* s) { return std::max(1, std::min(8, 256 / std::max(1, s->grid.ncell))); }
static int direct_blocks(const lg_system* s) { return s->grid.ncell * direct_split(s); }

// Atoms per block of the rho pass: the narrowest chunk (64..512) that keeps
// the grid at or below ~2048 blocks, so that per-block work stays
// substantial while the partials stay few; the tuning knob overrides.
static int rho_chunk(const lg_system* s) {
    if (s->tune.rho_width > 0) return s->tune.rho_width;
    int nub = blocks_for(s->nunits, 64);
    int w = 64;
    while (w < 512 && size_t(nub) * size_t(blocks_for(s->ld, w)) > 2048) w *= 2;
    return w;
}

// The matrix-core rho form when asked for (rho_kind 2; the row form is the
// default: it measured as fast or faster on the bench systems), kmax <= 16
// and its LDS (the larger of its own and the Verlet build's, which share a
// launch) within the device limit.
static bool rho_use_mfma(const lg_system* s, bool mixed) {
    if (s->mfma_nt <= 0 || s->tune.rho_kind != 2) return false;
    const RhoShape sh = rho_mfma_shape(s, mixed);
    return rho_mfma_bytes(sh.width, s->coul_kmax, mixed) <= s->ctx->max_shared;
}

// The next force evaluation rebuilds the list whatever the displacements.
static int force_rebuild(lg_system* s) {
    k_set_flag<<<1, 1, 0, S(s)>>>(s->flags.p, 1);
    LG_HIP(hipGetLastError());
    return LG_OK;
}

int launch_nlist(lg_system* s, bool check) {
    hipStream_t st = S(s);
    // profiled intervals only around launches (an empty event pair would
    // count the event packets' own dispatch as list time)
    long prof = -1;
    int n = int(s->n);
    int* flag = s->flags.p;
    const Grid& g = s->grid;
    const bool direct = direct_build(s);
    if (check) {
        prof = prof_start(s, LG_KERNEL_NLIST);
        double hs = 0.5 * s->skin;
        k_nlist_check<<<blocks_for(n, 256), 256, 0, st>>>(s->pos.p, s->ref.p, n, hs * hs, flag, cell_tags(s));
    }
    if (direct) {
        // the cell tags of the current positions are written (here or by
        // the MD first half-step); no cell list
        const float rl2 = (!s->periodic || s->rlist <= 0) ? INFINITY
                                                           : float(s->rlist * (1.0 + 1e-4) * s->rlist * (1.0 + 1e-4));
        DirectArgs da{s->pos.p, s->meta.p, s->cellid.p, n, g, s->box, rl2, s->cap,
                      s->nbr.p, s->nnbr.p, s->ref.p, s->sorted.p, flag, direct_split(s)};
        const bool nearest = s->periodic && (g.reach[0] == 0 || g.reach[1] == 0 || g.reach[2] == 0);
        prof_stop(s, prof);
        if (s->coul_kind == COUL_EWALD && s->nk > 0) {
            // launched by launch_forces together with the rho(k) partials
            s->direct_pending = true;
            s->cells_nearest = nearest;
            s->direct_args = da;
            return LG_OK;
        }
        long p2 = prof_start(s, LG_KERNEL_NLIST);
        if (nearest)
            k_nlist_direct<true><<<direct_blocks(s), kCellBlock, 0, st>>>(da);
        else
            k_nlist_direct<false><<<direct_blocks(s), kCellBlock, 0, st>>>(da);
        prof_stop(s, p2);
        LG_HIP(hipGetLastError());
        return LG_OK;
    }
    if (prof < 0) prof = prof_start(s, LG_KERNEL_NLIST);
    // small systems: the whole cell list in one workgroup
    size_t small_lds = size_t(8) * g.ncell + 4 + size_t(8) * n;
    if (g.ncell < 65536 && n <= 32768 && small_lds <= s->ctx->max_shared) {
        k_cell_small<<<1, 1024, small_lds, st>>>(s->pos.p, s->meta.p, n, g, s->box, s->cellid.p, s->start.p,
                                                 s->sorted.p, s->cpos.p, flag);
    } else {
        k_cell_count<<<blocks_for(n, 256), 256, 0, st>>>(s->pos.p, n, g, s->box, s->cellid.p, s->slot.p,
                                                          s->count.p, flag);
        k_cell_scan<<<1, 1024, 0, st>>>(s->count.p, s->start.p, g.ncell, flag);
        k_cell_scatter<<<blocks_for(n, 256), 256, 0, st>>>(s->cellid.p, s->slot.p, s->start.p, s->tmp.p, n,
                                                            flag);
        k_cell_rank<<<std::min(blocks_for(g.ncell, 4), kRankBlocks), 256, 0, st>>>(
            s->tmp.p, s->start.p, s->sorted.p, s->cpos.p, s->pos.p, s->meta.p, s->box, g.ncell, flag);
    }
    int nb = blocks_for(size_t(n) * s->tpa, 256);
    // FP32 candidate test with a 1e-4 relative margin on rlist
    float rl2f = (!s->periodic || s->rlist <= 0) ? INFINITY : float(s->rlist * (1.0 + 1e-4) * s->rlist * (1.0 + 1e-4));
    if (cell_build(s)) {
        CellsArgs ca{s->meta.p, s->cpos.p, s->sorted.p, s->start.p, s->count.p, g, s->box, rl2f, s->cap,
                     s->nbr.p, s->nnbr.p, s->pos.p, s->ref.p, flag};
        const bool nearest = s->periodic && (g.reach[0] == 0 || g.reach[1] == 0 || g.reach[2] == 0);
        prof_stop(s, prof);
        if (s->coul_kind == COUL_EWALD && s->nk > 0) {
            // launched by launch_forces together with the rho(k) partials
            s->cells_pending = true;
            s->cells_nearest = nearest;
            s->cells_args = ca;
            return LG_OK;
        }
        long p2 = prof_start(s, LG_KERNEL_NLIST);
        if (nearest)
            k_nlist_cells<true><<<g.ncell, kCellBlock, 0, st>>>(ca);
        else
            k_nlist_cells<false><<<g.ncell, kCellBlock, 0, st>>>(ca);
        prof_stop(s, p2);
        LG_HIP(hipGetLastError());
        return LG_OK;
    }
    switch (s->tpa) {
#define LG_BUILD_CASE(T)                                                                                   \
    case T:                                                                                                \
        k_nlist_build<T><<<nb, 256, 0, st>>>(s->cpos.p, s->sorted.p, s->cellid.p, s->start.p, s->count.p, g,  \
                                             s->box, rl2f, n, s->cap, s->nbr.p, s->nnbr.p, s->pos.p, s->ref.p,  \
                                             flag, s->meta.p);                                                \
        break;
        LG_BUILD_CASE(4)
        LG_BUILD_CASE(8)
        LG_BUILD_CASE(16)
        LG_BUILD_CASE(32)
        LG_BUILD_CASE(64)
#undef LG_BUILD_CASE
        default:
            return fail(LG_E_DEVICE, "invalid lanes-per-atom setting");
    }
    prof_stop(s, prof);
    LG_HIP(hipGetLastError());
    return LG_OK;
}

// The arguments of the intra-molecular terms (k_intra_terms, or blocks of
// another launch).
static IntraArgs intra_args(lg_system* s) {
    IntraArgs ia{};
    ia.pos = s->pos.p;
    ia.meta = s->meta.p;
    ia.nterms = s->nterms;
    ia.terms = s->terms.p;
    ia.table = s->table.p;
    ia.tabs = s->ptabs.p;
    ia.nkinds = int(s->nkinds);
    ia.pair_restr = s->pair_restr.p;
    ia.pair_s14 = s->pair_s14.p;
    ia.bparam = s->bparam.p;
    ia.aparam = s->aparam.p;
    ia.dparam = s->dparam.p;
    ia.coul = s->coul;
    ia.coul_restr = s->coul_restr;
    ia.coul_s14 = s->coul_s14;
    ia.box = s->box;
    ia.contrib = s->contrib.p;
    ia.epart = s->ep_mol_at ? s->ep_mol_at : s->epart_mol.p;
    return ia;
}

// phase_done: the MD first half-step has already written the phase table
// and the FP32 positions (frac) of the current positions.
int launch_forces(lg_system* s, bool energy, bool combine = true, bool phase_done = false) {
    hipStream_t st = S(s);
    int n = int(s->n);
    bool mixed = s->precision == LG_PRECISION_MIXED;
    // rho(k) partials first: their reduction runs as the last blocks of the
    // pair launch, and the k-space forces after it
    int nchunks = 0;
    const bool use_mfma = s->coul_kind == COUL_EWALD && s->nk > 0 && rho_use_mfma(s, mixed);
    // mixed form, forces only: the intra-molecular terms ride as the last
    // blocks of the direct-build list-and-rho launch (its 64 KB of LDS hold
    // it to 4 waves per SIMD, which the terms' 104 VGPRs do not lower)
    const bool intra_fused = mixed && !energy && !s->gen_pairs && s->nterms > 0 && s->coul_kind == COUL_EWALD &&
                             s->nk > 0 && !use_mfma && s->direct_pending;
    const IntraArgs fia = intra_fused ? intra_args(s) : IntraArgs{};
    const int fib = intra_fused ? blocks_for(s->nterms, kCellBlock) : 0;
    if (s->coul_kind == COUL_EWALD && s->nk > 0) {
        int kmax = s->coul_kmax;
        int nub = blocks_for(s->nunits, 64);
        int w = rho_chunk(s);
        nchunks = blocks_for(s->ld, w);
        unsigned grid = xcd_grid(nub, nchunks);
        // with a deferred Verlet build: one launch of k_cells_rho, the list
        // blocks first (cells rounded up to 8), then the rho blocks
        const bool direct = s->direct_pending;
        const bool cells = s->cells_pending || direct;
        const int ncb = cells ? ((direct ? direct_blocks(s) : s->grid.ncell) + 7) / 8 * 8 : 0;
        const CellsArgs& ca = s->cells_args;
        const DirectArgs& da = s->direct_args;
        const bool near = s->cells_nearest;
        if (use_mfma) {
            // the matrix-core rho shares its launch with the cell-list build
            // only; a direct build runs first on its own
            if (direct && near)
                launch_timed(s, LG_KERNEL_KSPACE_RHO, k_nlist_direct<true>, direct_blocks(s), kCellBlock, 0, da);
            else if (direct)
                launch_timed(s, LG_KERNEL_KSPACE_RHO, k_nlist_direct<false>, direct_blocks(s), kCellBlock, 0, da);
            const bool cells = s->cells_pending;
            const RhoShape sh = rho_mfma_shape(s, mixed);
            nchunks = sh.nchunks;
            const unsigned mgrid = xcd_grid(sh.ngroups, sh.nchunks);
            if (mixed) {
                if (!phase_done)
                    launch_timed(s, LG_KERNEL_KSPACE_RHO, k_phase32, blocks_for(size_t(3) * s->ld, 256), 256, 0, s->pos.p, n, s->ld, s->box, kmax,
                                                                                  s->phase_tab32.p, s->phase_q32.p);
                RhoMfmaArgs<float, float2> ra{s->pos.p, n, s->box, s->ld, kmax, sh.wlog, sh.cpb,
                                              s->mfma_nt, sh.ngroups, sh.nchunks, s->mcols.p, s->mkmap.p, s->nk,
                                              s->rho_partial.p};
                const size_t lds = rho_mfma_bytes(sh.width, kmax, true);
                if (!cells)
                    launch_timed(s, LG_KERNEL_KSPACE_RHO, k_rho_mfma<float, float2>, mgrid, 512, lds, ra);
                else if (near)
                    launch_timed(s, LG_KERNEL_KSPACE_RHO, k_cells_rho_mfma<true, float, float2>, ncb + mgrid, kCellBlock, lds, ca, ra, ncb);
                else
                    launch_timed(s, LG_KERNEL_KSPACE_RHO, k_cells_rho_mfma<false, float, float2>, ncb + mgrid, kCellBlock, lds, ca, ra, ncb);
            } else {
                if (!phase_done)
                    launch_timed(s, LG_KERNEL_KSPACE_RHO, k_phase, blocks_for(size_t(3) * s->ld, 256), 256, 0, s->pos.p, n, s->ld, s->box, kmax,
                                                                                s->phase_tab.p, s->phase_q.p);
                RhoMfmaArgs<double, double2> ra{s->pos.p, n, s->box, s->ld, kmax, sh.wlog, sh.cpb,
                                                s->mfma_nt, sh.ngroups, sh.nchunks, s->mcols.p, s->mkmap.p, s->nk,
                                                s->rho_partial.p};
                const size_t lds = rho_mfma_bytes(sh.width, kmax, false);
                if (!cells)
                    launch_timed(s, LG_KERNEL_KSPACE_RHO, k_rho_mfma<double, double2>, mgrid, 512, lds, ra);
                else if (near)
                    launch_timed(s, LG_KERNEL_KSPACE_RHO, k_cells_rho_mfma<true, double, double2>, ncb + mgrid, kCellBlock, lds, ca, ra, ncb);
                else
                    launch_timed(s, LG_KERNEL_KSPACE_RHO, k_cells_rho_mfma<false, double, double2>, ncb + mgrid, kCellBlock, lds, ca, ra, ncb);
            }
        } else if (mixed) {
            if (!phase_done)
                launch_timed(s, LG_KERNEL_KSPACE_RHO, k_phase32, blocks_for(size_t(3) * s->ld, 256), 256, 0, s->pos.p, n, s->ld, s->box, kmax,
                                                                              s->phase_tab32.p, s->phase_q32.p);
            RhoRowsArgs<float, float2> ra{s->phase_tab32.p, s->phase_q32.p, s->ld, kmax, w, s->units.p,
                                          s->nunits, s->rho_partial.p, nub, nchunks, s->tune.rho_kind != 3};
            if (!cells)
                launch_timed(s, LG_KERNEL_KSPACE_RHO, k_rho_rows<float, float2>, grid, 512, 0, ra);
            else if (direct && near) {
                if (fib > 0)
                    launch_timed(s, LG_KERNEL_KSPACE_RHO, k_direct_rho<true, float, float2, true>, ncb + grid + fib,
                                 kCellBlock, 0, da, ra, ncb, int(grid), fia);
                else
                    launch_timed(s, LG_KERNEL_KSPACE_RHO, k_direct_rho<true, float, float2, false>, ncb + grid,
                                 kCellBlock, 0, da, ra, ncb, int(grid), fia);
            }
            else if (direct) {
                if (fib > 0)
                    launch_timed(s, LG_KERNEL_KSPACE_RHO, k_direct_rho<false, float, float2, true>, ncb + grid + fib,
                                 kCellBlock, 0, da, ra, ncb, int(grid), fia);
                else
                    launch_timed(s, LG_KERNEL_KSPACE_RHO, k_direct_rho<false, float, float2, false>, ncb + grid,
                                 kCellBlock, 0, da, ra, ncb, int(grid), fia);
            }
            else if (near)
                launch_timed(s, LG_KERNEL_KSPACE_RHO, k_cells_rho<true, float, float2>, ncb + grid, kCellBlock, 0, ca, ra, ncb);
            else
                launch_timed(s, LG_KERNEL_KSPACE_RHO, k_cells_rho<false, float, float2>, ncb + grid, kCellBlock, 0, ca, ra, ncb);
        } else {
            if (!phase_done)
                launch_timed(s, LG_KERNEL_KSPACE_RHO, k_phase, blocks_for(size_t(3) * s->ld, 256), 256, 0, s->pos.p, n, s->ld, s->box, kmax,
                                                                            s->phase_tab.p, s->phase_q.p);
            RhoRowsArgs<double, double2> ra{s->phase_tab.p, s->phase_q.p, s->ld, kmax, w, s->units.p,
                                            s->nunits, s->rho_partial.p, nub, nchunks, s->tune.rho_kind != 3};
            if (!cells)
                launch_timed(s, LG_KERNEL_KSPACE_RHO, k_rho_rows<double, double2>, grid, 512, 0, ra);
            else if (direct && near)
                launch_timed(s, LG_KERNEL_KSPACE_RHO, k_direct_rho<true, double, double2>, ncb + grid, kCellBlock, 0, da, ra,
                             ncb, int(grid), IntraArgs{});
            else if (direct)
                launch_timed(s, LG_KERNEL_KSPACE_RHO, k_direct_rho<false, double, double2>, ncb + grid, kCellBlock, 0, da, ra,
                             ncb, int(grid), IntraArgs{});
            else if (near)
                launch_timed(s, LG_KERNEL_KSPACE_RHO, k_cells_rho<true, double, double2>, ncb + grid, kCellBlock, 0, ca, ra, ncb);
            else
                launch_timed(s, LG_KERNEL_KSPACE_RHO, k_cells_rho<false, double, double2>, ncb + grid, kCellBlock, 0, ca, ra, ncb);
        }
        s->cells_pending = false;
        s->direct_pending = false;
    }
    if (s->cells_pending) {  // deferred build without k-space (not expected): launch it alone
        if (s->cells_nearest)
            k_nlist_cells<true><<<s->grid.ncell, kCellBlock, 0, st>>>(s->cells_args);
        else
            k_nlist_cells<false><<<s->grid.ncell, kCellBlock, 0, st>>>(s->cells_args);
        s->cells_pending = false;
    }
    if (s->direct_pending) {
        if (s->cells_nearest)
            k_nlist_direct<true><<<direct_blocks(s), kCellBlock, 0, st>>>(s->direct_args);
        else
            k_nlist_direct<false><<<direct_blocks(s), kCellBlock, 0, st>>>(s->direct_args);
        s->direct_pending = false;
    }
    PairArgs pa{};
    pa.pos = s->pos.p;
    pa.meta = s->meta.p;
    pa.sorted = s->sorted.p;
    pa.nbr = s->nbr.p;
    pa.nnbr = s->nnbr.p;
    pa.cap = s->cap;
    pa.table = s->table.p;
    pa.tabs = s->ptabs.p;
    pa.nkinds = int(s->nkinds);
    pa.coul = s->coul;
    pa.box = s->box;
    pa.n = n;
    pa.force = s->force.p;
    pa.epart = s->ep_pairs_at ? s->ep_pairs_at : s->epart_pairs.p;
    pa.rebuild = s->flags.p;
    pa.ecx64 = s->ecx_ok ? s->ecx64.p : nullptr;
    pa.ecx32 = s->ecx_ok ? s->ecx32.p : nullptr;
    pa.ecx_inv_h = s->ecx_inv_h;
    pa.hf32 = s->hf_ok ? s->hf32.p : nullptr;
    pa.hf_inv_h = s->hf_inv_h;
    pa.frac = s->frac.p;
    if (mixed && !s->gen_pairs && !phase_done) {
        // the MD half-steps write the FP32 positions beside the new ones;
        // any other caller converts the current positions here
        launch_timed(s, LG_KERNEL_PAIRS, k_frac, blocks_for(n, 256), 256, 0, s->pos.p, n, s->box, s->frac_origin,
                     s->frac.p);
    }
    // the intra-molecular terms run as the last blocks of the pair launch
    // (FP64 form), of the list-and-rho launch (mixed form, forces only,
    // direct build: intra_fused) or in a launch of their own
    IntraArgs& ia = pa.intra;
    ia = intra_args(s);
    const int iblocks = s->nterms > 0 ? blocks_for(s->nterms, 256) : 0;
    // rho(k) reduction blocks
    const bool kspace = s->coul_kind == COUL_EWALD && s->nk > 0;
    RhoReduceArgs& rr = pa.rr;
    rr = RhoReduceArgs{};
    int rblocks = 0;
    if (kspace) {
        rr.partial = s->rho_partial.p;
        rr.nchunks = nchunks;
        rr.nslot = use_mfma ? s->nk : kRhoUnitK * s->nunits;
        rr.kslot = use_mfma ? s->kslot_id.p : s->kslot.p;
        rr.nk = s->nk;
        rr.rho = s->rho.p;
        rr.rec = mixed ? nullptr : s->krec.p;
        rr.rec32 = mixed ? s->krec32.p : nullptr;
        rr.efac = s->efac.p;
        rr.vfac = s->vfac.p;
        rr.epart = s->ep_k_at ? s->ep_k_at : s->epart_k.p;
        rblocks = blocks_for(s->nk, kRhoReduceK);
    }
    // the mixed instance leaves the intra-molecular terms to a launch of
    // their own (see k_pairs_mixed)
    const bool intra_apart = mixed && !s->gen_pairs;
    if (intra_apart && iblocks > 0 && !intra_fused) {
        if (energy)
            launch_timed(s, LG_KERNEL_MOLECULES, k_intra_terms<true>, iblocks, 256, 0, ia);
        else
            launch_timed(s, LG_KERNEL_MOLECULES, k_intra_terms<false>, iblocks, 256, 0, ia);
    }
    const int pblocks = pair_blocks_of(size_t(n), s->tpa);
    pa.pair_blocks = pblocks;
    pa.intra_blocks = intra_apart ? 0 : iblocks;
    const int nblocks = pblocks + pa.intra_blocks + rblocks;
    const int ck = coulomb_variant(s->coul);
#define LG_PAIRS_CK(T, CK)                                                      \
    if (ck == CK) {                                                             \
        if (mixed && energy)                                                    \
            launch_timed(s, LG_KERNEL_PAIRS, k_pairs_mixed<true, T, CK>, nblocks, kPairBlock, 0, pa);     \
        else if (mixed)                                                         \
            launch_timed(s, LG_KERNEL_PAIRS, k_pairs_mixed<false, T, CK>, nblocks, kPairBlock, 0, pa);    \
        else if (energy)                                                        \
            launch_timed(s, LG_KERNEL_PAIRS, k_pairs<true, T, CK>, nblocks, kPairBlock, 0, pa);           \
        else                                                                    \
            launch_timed(s, LG_KERNEL_PAIRS, k_pairs<false, T, CK>, nblocks, kPairBlock, 0, pa);          \
    }
#define LG_PAIRS_CASE(T)                 \
    case T:                              \
        LG_PAIRS_CK(T, CK_NONE)          \
        LG_PAIRS_CK(T, CK_EWALD)         \
        LG_PAIRS_CK(T, CK_EWALD_ERF)     \
        LG_PAIRS_CK(T, CK_WOLF)          \
        break;
    if (s->gen_pairs) {
        // any pair potential kind or a table: the general FP64 instance, in
        // both forms
        if (energy) {
            if (ck == CK_NONE) launch_timed(s, LG_KERNEL_PAIRS, k_pairs<true, 16, CK_NONE, true>, nblocks, kPairBlock, 0, pa);
            if (ck == CK_EWALD) launch_timed(s, LG_KERNEL_PAIRS, k_pairs<true, 16, CK_EWALD, true>, nblocks, kPairBlock, 0, pa);
            if (ck == CK_EWALD_ERF) launch_timed(s, LG_KERNEL_PAIRS, k_pairs<true, 16, CK_EWALD_ERF, true>, nblocks, kPairBlock, 0, pa);
            if (ck == CK_WOLF) launch_timed(s, LG_KERNEL_PAIRS, k_pairs<true, 16, CK_WOLF, true>, nblocks, kPairBlock, 0, pa);
        } else {
            if (ck == CK_NONE) launch_timed(s, LG_KERNEL_PAIRS, k_pairs<false, 16, CK_NONE, true>, nblocks, kPairBlock, 0, pa);
            if (ck == CK_EWALD) launch_timed(s, LG_KERNEL_PAIRS, k_pairs<false, 16, CK_EWALD, true>, nblocks, kPairBlock, 0, pa);
            if (ck == CK_EWALD_ERF) launch_timed(s, LG_KERNEL_PAIRS, k_pairs<false, 16, CK_EWALD_ERF, true>, nblocks, kPairBlock, 0, pa);
            if (ck == CK_WOLF) launch_timed(s, LG_KERNEL_PAIRS, k_pairs<false, 16, CK_WOLF, true>, nblocks, kPairBlock, 0, pa);
        }
    } else switch (s->tpa) {
        LG_PAIRS_CASE(4)
        LG_PAIRS_CASE(8)
        LG_PAIRS_CASE(16)
        LG_PAIRS_CASE(32)
        LG_PAIRS_CASE(64)
        default:
            return fail(LG_E_DEVICE, "invalid lanes-per-atom setting");
    }
#undef LG_PAIRS_CASE
#undef LG_PAIRS_CK
    if (s->coul_kind == COUL_EWALD && s->nk > 0) {
        int kmax = s->coul_kmax;
        constexpr int kApl = 1;  // atoms per lane (2 measured slower at 4k-9k atoms)
        int nab = blocks_for(n, 64 * kApl);
        unsigned fgrid = xcd_grid(s->ksplit, nab);
        if (mixed)
            launch_timed(s, LG_KERNEL_KSPACE_FORCE, k_kspace_force<float, float2, KRec32, kApl>, fgrid, 256, 0, 
                s->phase_tab32.p, s->ld, n, kmax, s->rows.p, s->nrows, s->krec32.p, s->kg[0], s->kg[1], s->kg[2],
                s->ksplit, nab, s->fk.p);
        else
            launch_timed(s, LG_KERNEL_KSPACE_FORCE, k_kspace_force<double, double2, KRec, kApl>, fgrid, 256, 0, 
                s->phase_tab.p, s->ld, n, kmax, s->rows.p, s->nrows, s->krec.p, s->kg[0], s->kg[1], s->kg[2],
                s->ksplit, nab, s->fk.p);
    }
    if (combine && (kspace || s->nterms > 0)) {
        // timed with the k-space forces: it adds their partials
        long pc = prof_start(s, kspace ? LG_KERNEL_KSPACE_FORCE : LG_KERNEL_MOLECULES);
        k_force_combine<<<blocks_for(n, 256), 256, 0, st>>>(s->force.p, s->contrib.p,
                                                           s->nterms > 0 ? s->slot_start.p : nullptr, s->fk.p,
                                                           kspace ? s->ksplit : 0, s->pos.p, n);
        prof_stop(s, pc);
    }
    LG_HIP(hipGetLastError());
    return LG_OK;
}

// Reads the list-overflow flag (flags[1]) and, when `nonfinite` is given, the
// non-finite-state flag of the integrator (flags[2]).
int check_overflow(lg_system* s, bool* overflowed, bool* nonfinite = nullptr) {
    int fl[2] = {0, 0};
    LG_HIP(hipMemcpyAsync(fl, s->flags.p + 1, sizeof(fl), hipMemcpyDeviceToHost, S(s)));
    LG_HIP(hipStreamSynchronize(S(s)));
    const int ov = fl[0];
    *overflowed = ov > 0;
    if (nonfinite) *nonfinite = fl[1] != 0;
    if (ov > 0) {
        s->cap = int(std::min<size_t>(s->n, size_t(ov) + size_t(ov) / 4 + 32));
        LG_TRY(s->nbr.ensure(size_t(s->cap) * s->n));
        LG_TRY(force_rebuild(s));
        int zero = 0;
        LG_HIP(hipMemcpyAsync(s->flags.p + 1, &zero, sizeof(int), hipMemcpyHostToDevice, S(s)));
        LG_HIP(hipStreamSynchronize(S(s)));
    }
    return LG_OK;
}

// Forces (and optionally energy partials) for the current positions.
int compute(lg_system* s, bool energy) {
    LG_TRY(prepare(s));
    for (int attempt = 0; attempt < 4; attempt++) {
        LG_TRY(launch_nlist(s, true));
        LG_TRY(launch_forces(s, energy));
        bool ov = false;
        LG_TRY(check_overflow(s, &ov));
        if (!ov) return LG_OK;
    }
    return fail(LG_E_DEVICE, "neighbour list overflow could not be resolved");
}

struct Energies {
    double pairs = 0, coul_real = 0, bonds = 0, angles = 0, dihedrals = 0, kspace = 0;
    double w[6] = {0, 0, 0, 0, 0, 0};      // pairs + coulomb real + bonds
    double wk[6] = {0, 0, 0, 0, 0, 0};     // k-space
};

int gather_energies(lg_system* s, Energies& E) {
    hipStream_t st = S(s);
    int nbp = pair_blocks_of(s->n, s->tpa);
    std::vector<double> hp(size_t(nbp) * 8);
    LG_HIP(hipMemcpyAsync(hp.data(), s->epart_pairs.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    std::vector<double> hm;
    int nbm = s->nterms > 0 ? blocks_for(s->nterms, 256) : 0;
    if (nbm) {
        hm.resize(size_t(nbm) * 11);
        LG_HIP(hipMemcpyAsync(hm.data(), s->epart_mol.p, hm.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    std::vector<double> hk;
    int nbk = (s->coul_kind == COUL_EWALD && s->nk > 0) ? blocks_for(s->nk, kRhoReduceK) : 0;
    if (nbk) {
        hk.resize(size_t(nbk) * 7);
        LG_HIP(hipMemcpyAsync(hk.data(), s->epart_k.p, hk.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    LG_HIP(hipStreamSynchronize(st));
    for (int b = 0; b < nbp; b++) {
        E.pairs += hp[b * 8 + 0];
        E.coul_real += hp[b * 8 + 1];
        for (int q = 0; q < 6; q++) E.w[q] += hp[b * 8 + 2 + q];
    }
    for (int b = 0; b < nbm; b++) {
        E.pairs += hm[b * 11 + 0];
        E.coul_real += hm[b * 11 + 1];
        E.bonds += hm[b * 11 + 2];
        E.angles += hm[b * 11 + 3];
        E.dihedrals += hm[b * 11 + 4];
        for (int q = 0; q < 6; q++) E.w[q] += hm[b * 11 + 5 + q];
    }
    for (int b = 0; b < nbk; b++) {
        E.kspace += hk[b * 7 + 0];
        for (int q = 0; q < 6; q++) E.wk[q] += hk[b * 7 + 1 + q];
    }
    E.kspace /= kFourPiEps0;
    for (int q = 0; q < 6; q++) E.wk[q] /= kFourPiEps0;
    return LG_OK;
}

// composition.rs:123-128 then energy.rs:63-79 / compute.rs:218-228
void tail_terms(lg_system* s, double* e_tail, double* w_tail_diag) {
    *e_tail = 0.0;
    *w_tail_diag = 0.0;
    if (!s->periodic) return;
    std::vector<double> count(s->nkinds, 0.0);
    for (size_t i = 0; i < s->n; i++) count[s->kind[i]] += 1.0;
    double volume = s->L[0] * s->L[1] * s->L[2];
    for (uint32_t a = 0; a < s->nkinds; a++) {
        if (count[a] == 0) continue;
        for (uint32_t b = 0; b < s->nkinds; b++) {
            if (count[b] == 0) continue;
            double tpd = 2.0 * kPi * count[a] * count[b] / volume;
            auto it = s->pairs.find({std::min(a, b), std::max(a, b)});
            if (it == s->pairs.end()) continue;
            *e_tail += tpd * host_tail_energy(it->second);
            *w_tail_diag += tpd * (host_tail_virial(it->second) * (1.0 / 3.0));
        }
    }
}

double coulomb_self(lg_system* s) {
    if (s->coul_kind == COUL_EWALD) {
        double q2 = 0.0;
        for (size_t i = 0; i < s->n; i++) q2 += s->charge[i] * s->charge[i];
        return -s->coul_alpha / std::sqrt(kPi) * q2 / kFourPiEps0;
    }
    if (s->coul_kind == COUL_WOLF) {
        double e = 0.0;
        const CoulParam& c = s->coul;
        for (size_t i = 0; i < s->n; i++) {
            double q = s->charge[i];
            if (q == 0.0) continue;
            e -= q * q * 0.5 * (c.e_const + c.alpha * kFrac2SqrtPi) / kFourPiEps0;
        }
        return e;
    }
    return 0.0;
}

int push_vec3(DevBuf<double4>& buf, size_t n, const double* xyz, const std::vector<double>& w, hipStream_t st) {
    std::vector<double4> h(n);
    for (size_t i = 0; i < n; i++) h[i] = make_double4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], w[i]);
    LG_HIP(hipMemcpyAsync(buf.p, h.data(), n * sizeof(double4), hipMemcpyHostToDevice, st));
    LG_HIP(hipStreamSynchronize(st));
    return LG_OK;
}

int pull_vec3(DevBuf<double4>& buf, size_t n, double* xyz, hipStream_t st, bool accumulate = false) {
    std::vector<double4> h(n);
    LG_HIP(hipMemcpyAsync(h.data(), buf.p, n * sizeof(double4), hipMemcpyDeviceToHost, st));
    LG_HIP(hipStreamSynchronize(st));
    for (size_t i = 0; i < n; i++) {
        if (accumulate) {
            xyz[3 * i] += h[i].x;
            xyz[3 * i + 1] += h[i].y;
            xyz[3 * i + 2] += h[i].z;
        } else {
            xyz[3 * i] = h[i].x;
            xyz[3 * i + 1] = h[i].y;
            xyz[3 * i + 2] = h[i].z;
        }
    }
    return LG_OK;
}

// replace the .w lane of a double4 buffer
__global__ void k_set_w(double4* __restrict__ b, const double* __restrict__ w, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i].w = w[i];
}

int set_w(lg_system* s, DevBuf<double4>& b, const std::vector<double>& w) {
    DevBuf<double> tmp;
    LG_TRY(tmp.ensure(s->n));
    LG_HIP(hipMemcpyAsync(tmp.p, w.data(), s->n * sizeof(double), hipMemcpyHostToDevice, S(s)));
    k_set_w<<<blocks_for(s->n, 256), 256, 0, S(s)>>>(b.p, tmp.p, int(s->n));
    LG_HIP(hipGetLastError());
    LG_HIP(hipStreamSynchronize(S(s)));
    tmp.release();
    return LG_OK;
}

// ---- Monte Carlo ---------------------------------------------------------------------
// Sets the neighbour-list rebuild flag after positions changed on the device.
int flag_rebuild(lg_system* s) {
    if (s->nlist_dirty || !s->flags.p) return LG_OK;
    return force_rebuild(s);
}

// EnergyCache::energy: the pair and coulomb terms live on the device (updated
// by every accepted move), the others are unchanged by translations.
int mc_total(lg_system* s, double* e) {
    double pc[2];
    LG_HIP(hipMemcpyAsync(pc, s->mc_out.p, sizeof(pc), hipMemcpyDeviceToHost, S(s)));
    LG_HIP(hipStreamSynchronize(S(s)));
    *e = pc[0] + s->c_tail + s->c_bonds + s->c_angles + pc[1];
    return LG_OK;
}

bool mc_ewald(const lg_system* s) { return s->coul_kind == COUL_EWALD && s->nk > 0; }

size_t mc_phase_bytes(const lg_system* s) {
    if (!mc_ewald(s)) return 0;
    return size_t(std::min(s->max_mol, kMcMaxMol)) * 2 * size_t(phase_rows(s->coul_kmax)) * sizeof(double2);
}

// EnergyCache::init (cache.rs:75-98): the energy terms of the current
// configuration and, with Ewald, rho(k) (ewald.rs:222).
int mc_init(lg_system* s) {
    lg_energy_terms t{};
    LG_TRY(lg_potential_energy(s, &t));
    s->c_pairs = t.pairs;
    s->c_tail = t.pairs_tail;
    s->c_bonds = t.bonds;
    s->c_angles = t.angles + t.dihedrals;
    s->c_coul = t.coulomb;
    hipStream_t st = S(s);
    if (mc_ewald(s)) {
        LG_TRY(s->mc_rho.ensure(s->nk));
        LG_TRY(s->mc_drho.ensure(s->nk));
        LG_HIP(hipMemcpyAsync(s->mc_rho.p, s->rho.p, s->nk * sizeof(double2), hipMemcpyDeviceToDevice, st));
    }
    LG_TRY(s->mc_newpos.ensure(kMcMaxMol));
    LG_TRY(s->mc_out.ensure(8));
    double pc[2] = {s->c_pairs, s->c_coul};
    LG_HIP(hipMemcpyAsync(s->mc_out.p, pc, sizeof(pc), hipMemcpyHostToDevice, st));
    LG_HIP(hipStreamSynchronize(st));
    s->mc_ready = true;
    s->mc_staged = false;
    s->mc_cells_valid = false;
    return LG_OK;
}

// Cell list of the device chain: cells of edge >= the largest cutoff, at
// least 4 per axis (else the all-atoms loop); capacity 4x the mean
// occupancy + 32, or every atom when `full`.
int mc_prepare_cells(lg_system* s, bool full) {
    s->mc_ncell = 0;
    s->mc_cells_valid = false;
    double rc = max_cutoff(s);
    if (!s->periodic || rc <= 0.0) return LG_OK;
    int nc[3];
    for (int d = 0; d < 3; d++) {
        nc[d] = int(std::floor(s->L[d] / (rc * (1.0 + 1e-12))));
        if (nc[d] < 4) return LG_OK;
    }
    // the counts live in LDS: coarsen the largest axis until they fit
    while (nc[0] * nc[1] * nc[2] > kMcMaxCells) {
        int d = nc[0] >= nc[1] && nc[0] >= nc[2] ? 0 : (nc[1] >= nc[2] ? 1 : 2);
        nc[d]--;
    }
    int ncell = nc[0] * nc[1] * nc[2];
    size_t mean = (s->n + ncell - 1) / ncell;
    size_t cap = full ? s->n : std::min<size_t>(s->n, 4 * mean + 32);
    cap = std::max(cap, size_t(std::max(s->mc_ccap, 0)));  // keep a capacity grown earlier
    for (int d = 0; d < 3; d++) {
        s->mc_nc[d] = nc[d];
        s->mc_csize[d] = s->L[d] / nc[d];
    }
    LG_TRY(s->mc_cell_pos.ensure(size_t(ncell) * cap));
    LG_TRY(s->mc_cell_ids.ensure(size_t(ncell) * cap));
    LG_TRY(s->mc_cell_count.ensure(ncell));
    LG_TRY(s->mc_atom_cell.ensure(s->n));
    s->mc_ncell = ncell;
    s->mc_ccap = int(cap);
    return LG_OK;
}

McChain mc_chain(lg_system* s, bool cells) {
    McChain c{};
    c.pos = s->pos.p;
    c.vel = s->vel.p;
    c.meta = s->meta.p;
    c.mol_start = s->d_mol_start.p;
    c.n = int(s->n);
    c.nmol = s->nmol;
    c.box = s->box;
    c.table = s->table.p;
    c.tabs = s->ptabs.p;
    c.nkinds = int(s->nkinds);
    c.coul = s->coul;
    if (mc_ewald(s)) {
        c.kmax = s->coul_kmax;
        c.nk = s->nk;
        c.kidx = s->kidx.p;
        c.efac = s->efac.p;
        c.rho = s->mc_rho.p;
        c.drho = s->mc_drho.p;
    }
    if (cells && s->mc_ncell > 0) {
        for (int d = 0; d < 3; d++) {
            c.nc[d] = s->mc_nc[d];
            c.csize[d] = s->mc_csize[d];
            c.inv_csize[d] = 1.0 / s->mc_csize[d];
        }
        c.ncell = s->mc_ncell;
        c.ccap = s->mc_ccap;
        c.cell_pos = s->mc_cell_pos.p;
        c.cell_ids = s->mc_cell_ids.p;
        c.cell_count = s->mc_cell_count.p;
        c.atom_cell = s->mc_atom_cell.p;
    }
    double rc = max_cutoff(s);
    c.r2_filter = rc * rc * (1.0 + 1e-6);
    c.wrap = 1;
    c.out = s->mc_out.p;
    return c;
}

}  // namespace

// ===================================================================================
// C ABI
// ===================================================================================
extern "C" {

const char* lg_last_error(void) { return g_error.c_str(); }
const char* lg_version(void) { return "lumol_gpu 0.1 gfx950"; }

int lg_context_create(int device, lg_context** out) {
    if (!out) return fail(LG_E_INVALID, "null output pointer");
    int count = 0;
    LG_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(LG_E_INVALID, "device index out of range");
    LG_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    LG_HIP(hipGetDeviceProperties(&prop, device));
    double* sp = nullptr;
    LG_TRY(scratch_of_device(device, &sp));
    lg_context* c = new lg_context();
    c->device = device;
    c->scratch = sp;
    c->max_shared = prop.sharedMemPerBlock;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(LG_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = c;
    return LG_OK;
}

int lg_context_destroy(lg_context* ctx) {
    if (!ctx) return LG_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    ctx->mc_desc.release();
    ctx->mc_cnt.release();
    delete ctx;
    return LG_OK;
}

int lg_system_create(lg_context* ctx, size_t natoms, lg_system** out) {
    if (!ctx || !out) return fail(LG_E_INVALID, "null argument");
    if (natoms == 0) return fail(LG_E_INVALID, "invalid number of atoms");
    // Verlet-list entries hold the atom index in 24 bits (nbr_entry)
    if (natoms > (size_t(1) << 24)) return fail(LG_E_UNSUPPORTED, "more than 2^24 atoms in one system");
    LG_HIP(hipSetDevice(ctx->device));
    lg_system* s = new lg_system();
    s->ctx = ctx;
    s->n = natoms;
    s->mass.assign(natoms, 1.0);
    s->charge.assign(natoms, 0.0);
    s->kind.assign(natoms, 0);
    int r;
    if ((r = s->pos.ensure(natoms)) || (r = s->vel.ensure(natoms)) || (r = s->acc.ensure(natoms)) ||
        (r = s->force.ensure(natoms)) || (r = s->lambda.ensure(1)) ||
        (r = s->ke_part.ensure(size_t(blocks_for(natoms, 64))))) {
        delete s;
        return r;
    }
    LG_HIP(hipMemsetAsync(s->pos.p, 0, natoms * sizeof(double4), ctx->stream));
    LG_HIP(hipMemsetAsync(s->acc.p, 0, natoms * sizeof(double4), ctx->stream));
    LG_HIP(hipMemsetAsync(s->force.p, 0, natoms * sizeof(double4), ctx->stream));
    std::vector<double4> v(natoms, make_double4(0, 0, 0, 1.0));
    LG_HIP(hipMemcpyAsync(s->vel.p, v.data(), natoms * sizeof(double4), hipMemcpyHostToDevice, ctx->stream));
    LG_HIP(hipStreamSynchronize(ctx->stream));
    *out = s;
    return LG_OK;
}

int lg_system_destroy(lg_system* s) {
    if (!s) return LG_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    DevBuf<int>* ibufs[] = {&s->cellid, &s->slot, &s->count, &s->start, &s->tmp, &s->sorted, &s->nbr, &s->nnbr,

                            &s->flags, &s->pair_restr, &s->slot_start, &s->d_mol_start,
                            &s->mc_cell_count, &s->mc_atom_cell, &s->kslot, &s->mcols, &s->mkmap, &s->kslot_id};
    for (auto* b : ibufs) b->release();
    s->units.release();
    s->ecx64.release();
    s->ecx32.release();
    s->hf32.release();
    s->phase_tab.release();
    s->phase_q.release();
    s->phase_tab32.release();
    s->phase_q32.release();
    s->krec32.release();
    s->krec.release();
    DevBuf<double4>* d4[] = {&s->pos, &s->vel, &s->acc, &s->force, &s->ref, &s->fk,
                             &s->snap_pos, &s->snap_vel, &s->snap_acc};
    for (auto* b : d4) b->release();
    DevBuf<double>* d1[] = {&s->pair_s14, &s->efac, &s->vfac, &s->epart_pairs, &s->epart_mol, &s->epart_k,
                            &s->ke_part, &s->lambda};
    for (auto* b : d1) b->release();
    for (auto e : s->prof_events) (void)hipEventDestroy(e);
    if (s->md_exec) (void)hipGraphExecDestroy(s->md_exec);
    s->md_exec = nullptr;
    s->meta.release();
    s->table.release();
    s->terms.release();
    s->contrib.release();
    s->mc_rho.release();
    s->mc_drho.release();
    s->mc_newpos.release();
    s->mc_out.release();
    s->mc_cell_pos.release();
    s->cpos.release();
    s->frac.release();
    s->mc_cell_ids.release();
    s->bparam.release();
    s->aparam.release();
    s->dparam.release();
    s->kidx.release();
    s->rows.release();
    s->rho.release();
    s->rho_partial.release();
    delete s;
    return LG_OK;
}

int lg_system_size(const lg_system* s, size_t* n) {
    if (!s || !n) return fail(LG_E_INVALID, "null argument");
    *n = s->n;
    return LG_OK;
}

int lg_system_set_cell(lg_system* s, double a, double b, double c) {
    if (!s) return fail(LG_E_INVALID, "null system");
    if (a == 0.0 && b == 0.0 && c == 0.0) {
        s->periodic = false;
    } else {
        if (!(a > 0.0 && b > 0.0 && c > 0.0) || !std::isfinite(a) || !std::isfinite(b) || !std::isfinite(c))
            return fail(LG_E_INVALID, "Cell lengths must be positive");
        s->periodic = true;
    }
    s->L[0] = a;
    s->L[1] = b;
    s->L[2] = c;
    s->kspace_dirty = true;
    s->nlist_dirty = true;
    return LG_OK;
}

int lg_system_set_positions(lg_system* s, const double* xyz) {
    if (!s || !xyz) return fail(LG_E_INVALID, "null argument");
    for (size_t i = 0; i < 3 * s->n; i++)
        if (!std::isfinite(xyz[i])) return fail(LG_E_NONFINITE, "non-finite position");
    LG_HIP(hipSetDevice(s->ctx->device));
    LG_TRY(push_vec3(s->pos, s->n, xyz, s->charge, S(s)));
    if (s->n > 0) s->frac_origin = make_double3(xyz[0], xyz[1], xyz[2]);
    s->mc_ready = s->mc_staged = s->mc_cells_valid = false;
    if (!s->nlist_dirty && s->flags.p) LG_TRY(force_rebuild(s));
    return LG_OK;
}

int lg_system_get_positions(lg_system* s, double* xyz) {
    if (!s || !xyz) return fail(LG_E_INVALID, "null argument");
    LG_HIP(hipSetDevice(s->ctx->device));
    return pull_vec3(s->pos, s->n, xyz, S(s));
}

int lg_system_set_velocities(lg_system* s, const double* v) {
    if (!s || !v) return fail(LG_E_INVALID, "null argument");
    LG_HIP(hipSetDevice(s->ctx->device));
    return push_vec3(s->vel, s->n, v, s->mass, S(s));
}

int lg_system_get_velocities(lg_system* s, double* v) {
    if (!s || !v) return fail(LG_E_INVALID, "null argument");
    LG_HIP(hipSetDevice(s->ctx->device));
    return pull_vec3(s->vel, s->n, v, S(s));
}

int lg_system_set_masses(lg_system* s, const double* m) {
    if (!s || !m) return fail(LG_E_INVALID, "null argument");
    for (size_t i = 0; i < s->n; i++)
        if (!(m[i] > 0.0) || !std::isfinite(m[i])) return fail(LG_E_INVALID, "masses must be positive");
    s->mass.assign(m, m + s->n);
    LG_HIP(hipSetDevice(s->ctx->device));
    return set_w(s, s->vel, s->mass);
}

int lg_system_set_charges(lg_system* s, const double* q) {
    if (!s || !q) return fail(LG_E_INVALID, "null argument");
    for (size_t i = 0; i < s->n; i++)
        if (!std::isfinite(q[i])) return fail(LG_E_NONFINITE, "non-finite charge");
    s->charge.assign(q, q + s->n);
    s->mc_ready = s->mc_staged = false;
    LG_HIP(hipSetDevice(s->ctx->device));
    return set_w(s, s->pos, s->charge);
}

int lg_system_set_kinds(lg_system* s, const uint32_t* kinds, uint32_t nkinds) {
    if (!s || !kinds || nkinds == 0) return fail(LG_E_INVALID, "null argument");
    if (nkinds > uint32_t(kMaxKinds)) return fail(LG_E_UNSUPPORTED, "more than 16 particle kinds");
    for (size_t i = 0; i < s->n; i++)
        if (kinds[i] >= nkinds) return fail(LG_E_INVALID, "kind index out of range");
    s->kind.assign(kinds, kinds + s->n);
    s->nkinds = nkinds;
    s->meta_dirty = true;
    s->params_dirty = true;
    s->topo_dirty = true;
    return LG_OK;
}

int lg_system_set_bonds(lg_system* s, size_t nbonds, const uint32_t* ij) {
    if (!s || (nbonds && !ij)) return fail(LG_E_INVALID, "null argument");
    std::vector<std::pair<uint32_t, uint32_t>> b;
    for (size_t k = 0; k < nbonds; k++) {
        uint32_t i = ij[2 * k], j = ij[2 * k + 1];
        if (i >= s->n || j >= s->n || i == j) return fail(LG_E_INVALID, "invalid bond");
        b.push_back({std::min(i, j), std::max(i, j)});
    }
    std::sort(b.begin(), b.end());
    b.erase(std::unique(b.begin(), b.end()), b.end());
    std::swap(s->bonds, b);
    s->topo_dirty = true;
    LG_HIP(hipSetDevice(s->ctx->device));
    std::vector<std::tuple<int, int, int>> ip;
    int r = build_topology(s, ip);  // validate now, as System::add_bond does
    if (r != LG_OK) {
        std::swap(s->bonds, b);
        s->topo_dirty = true;
        return r;
    }
    s->nlist_dirty = true;
    return LG_OK;
}

int lg_system_molecule_count(const lg_system* s, size_t* nmol) {
    if (!s || !nmol) return fail(LG_E_INVALID, "null argument");
    if (s->topo_dirty) {
        lg_system* ms = const_cast<lg_system*>(s);
        std::vector<std::tuple<int, int, int>> ip;
        LG_TRY(build_topology(ms, ip));
    }
    *nmol = size_t(s->nmol);
    return LG_OK;
}

int lg_system_set_precision(lg_system* s, int precision) {
    if (!s) return fail(LG_E_INVALID, "null system");
    if (precision != LG_PRECISION_FP64 && precision != LG_PRECISION_MIXED)
        return fail(LG_E_INVALID, "unknown precision");
    s->precision = precision;
    return LG_OK;
}

int lg_system_set_tuning(lg_system* s, const lg_tuning* t) {
    if (!s || !t) return fail(LG_E_INVALID, "null argument");
    int l = t->pair_lanes;
    if (l != 0 && l != 4 && l != 8 && l != 16 && l != 32 && l != 64)
        return fail(LG_E_INVALID, "pair_lanes must be 0, 4, 8, 16, 32 or 64");
    if (t->kspace_split < 0 || t->kspace_split > 32) return fail(LG_E_INVALID, "kspace_split must be in 0..32");
    // any multiple of 64 up to 512: the row kernel's eight waves then each
    // take a multiple of 8 atoms, and the partial-sum buffer is sized for
    // the narrowest width
    if (t->rho_width < 0 || t->rho_width > 512 || t->rho_width % 64 != 0)
        return fail(LG_E_INVALID, "rho_width must be 0 or a multiple of 64 up to 512");
    if (t->nlist_build < 0 || t->nlist_build > 3) return fail(LG_E_INVALID, "nlist_build must be 0, 1, 2 or 3");
    if (t->md_graph < 0 || t->md_graph > 2) return fail(LG_E_INVALID, "md_graph must be 0, 1 or 2");
    if (t->rho_kind < 0 || t->rho_kind > 3) return fail(LG_E_INVALID, "rho_kind must be 0, 1, 2 or 3");
    s->tune = *t;
    s->kspace_dirty = s->nlist_dirty = true;
    return LG_OK;
}

int lg_system_get_tuning(lg_system* s, lg_tuning* t) {
    if (!s || !t) return fail(LG_E_INVALID, "null argument");
    LG_HIP(hipSetDevice(s->ctx->device));
    LG_TRY(prepare(s));
    *t = s->tune;
    t->pair_lanes = s->tpa;
    t->kspace_split = (s->coul_kind == COUL_EWALD && s->nk > 0) ? s->ksplit : 0;
    t->nlist_build = direct_build(s) ? 3 : (cell_build(s) ? 2 : 1);
    t->md_graph = s->tune.md_graph == 2 ? 2 : 1;
    t->rho_width = (s->coul_kind == COUL_EWALD && s->nk > 0) ? rho_chunk(s) : 0;
    t->rho_kind = (s->coul_kind == COUL_EWALD && s->nk > 0)
                      ? (rho_use_mfma(s, s->precision == LG_PRECISION_MIXED) ? 2 : (s->tune.rho_kind == 3 ? 3 : 1))
                      : 0;
    return LG_OK;
}

static int set_pair(lg_system* s, uint32_t ki, uint32_t kj, const PotSpec& v, double cutoff, int shifted, int tail,
                    int restriction, double scale14) {
    if (ki >= s->nkinds || kj >= s->nkinds) return fail(LG_E_INVALID, "kind index out of range");
    if (restriction < 0 || restriction > 6) return fail(LG_E_INVALID, "unknown restriction");
    if (!(cutoff >= 0.0) || !std::isfinite(cutoff)) return fail(LG_E_INVALID, "invalid cutoff");
    if (s->periodic) {
        double half = std::min(s->L[0], std::min(s->L[1], s->L[2])) / 2.0;
        if (cutoff > half)
            return fail(LG_E_INVALID, "The cutoff for this pair potential is bigger than half the cell length");
    }
    PairSpec p{v, cutoff, shifted, tail, restriction, scale14};
    s->pairs[{std::min(ki, kj), std::max(ki, kj)}] = p;
    s->params_dirty = true;
    s->nlist_dirty = true;
    return LG_OK;
}

int lg_pair_set(lg_system* s, uint32_t ki, uint32_t kj, int pot, const double* params, double cutoff, int shifted,
                int tail, int restriction, double scale14) {
    if (!s) return fail(LG_E_INVALID, "null system");
    PotSpec v;
    LG_TRY(make_pot(pot, params, ROLE_PAIR, v));
    return set_pair(s, ki, kj, v, cutoff, shifted, tail, restriction, scale14);
}

int lg_pair_set_table(lg_system* s, uint32_t ki, uint32_t kj, int pot, const double* params, uint64_t table_size,
                      double table_max, double cutoff, int shifted, int tail, int restriction, double scale14) {
    if (!s) return fail(LG_E_INVALID, "null system");
    PotSpec v;
    LG_TRY(make_pot(pot, params, ROLE_PAIR, v));
    if (table_size < 1 || table_size > (uint64_t(1) << 24))
        return fail(LG_E_INVALID, "table size must be in [1, 2^24]");
    if (!(table_max > 0.0) || !std::isfinite(table_max)) return fail(LG_E_INVALID, "invalid table maximum");
    v.tab_n = table_size;
    v.tab_max = table_max;
    return set_pair(s, ki, kj, v, cutoff, shifted, tail, restriction, scale14);
}

int lg_bond_set(lg_system* s, uint32_t ki, uint32_t kj, int pot, const double* params) {
    if (!s) return fail(LG_E_INVALID, "null argument");
    BondedSpec v;
    LG_TRY(make_pot(pot, params, ROLE_BOND, v));
    s->bond_pots[{std::min(ki, kj), std::max(ki, kj)}] = v;
    s->topo_dirty = true;
    return LG_OK;
}

int lg_angle_set(lg_system* s, uint32_t ki, uint32_t kj, uint32_t kk, int pot, const double* params) {
    if (!s) return fail(LG_E_INVALID, "null argument");
    BondedSpec v;
    LG_TRY(make_pot(pot, params, ROLE_ANGLE, v));
    s->angle_pots[{std::min(ki, kk), kj, std::max(ki, kk)}] = v;
    s->topo_dirty = true;
    return LG_OK;
}

int lg_dihedral_set(lg_system* s, uint32_t ki, uint32_t kj, uint32_t kk, uint32_t km, int pot,
                    const double* params) {
    if (!s) return fail(LG_E_INVALID, "null argument");
    BondedSpec v;
    LG_TRY(make_pot(pot, params, ROLE_DIHEDRAL, v));
    uint32_t ij = std::max(ki, kj), kmm = std::max(kk, km);
    std::tuple<uint32_t, uint32_t, uint32_t, uint32_t> key;
    if (ij == kmm)
        key = std::min(ki, kj) < std::min(kk, km) ? std::make_tuple(ki, kj, kk, km) : std::make_tuple(km, kk, kj, ki);
    else if (ij < kmm)
        key = std::make_tuple(ki, kj, kk, km);
    else
        key = std::make_tuple(km, kk, kj, ki);
    s->dihedral_pots[key] = v;
    s->topo_dirty = true;
    return LG_OK;
}

// System::set_coulomb_potential (system.rs:157-167): the cutoff must not
// exceed half the smallest cell length of a periodic cell.
static int check_coulomb_cutoff(const lg_system* s, double cutoff) {
    if (s->periodic && cutoff > 0.5 * std::min(s->L[0], std::min(s->L[1], s->L[2])))
        return fail(LG_E_INVALID, "The cutoff for the coulomb potential is bigger than half the cell length");
    return LG_OK;
}

int lg_coulomb_ewald(lg_system* s, double cutoff, int kmax, double alpha, int restriction) {
    if (!s) return fail(LG_E_INVALID, "null system");
    if (std::isnan(alpha)) alpha = kPi / cutoff;  // ewald.rs:279
    if (!(cutoff >= 0.0)) return fail(LG_E_INVALID, "the cutoff can not be negative in Ewald");
    LG_TRY(check_coulomb_cutoff(s, cutoff));
    if (alpha < 0.0) return fail(LG_E_INVALID, "alpha can not be negative in Ewald");
    if (kmax <= 0) return fail(LG_E_INVALID, "kmax can not be 0 in Ewald");
    if (restriction == LG_RESTRICT_SCALE14)
        return fail(LG_E_UNSUPPORTED, "Scaling restriction scheme using Ewald are not implemented");
    if (restriction < 0 || restriction > 6) return fail(LG_E_INVALID, "unknown restriction");
    s->coul_kind = COUL_EWALD;
    s->coul_rc = cutoff;
    s->coul_alpha = alpha;
    s->coul_kmax = kmax;
    s->coul_restr = restriction;
    s->coul_s14 = 1.0;
    s->params_dirty = s->kspace_dirty = s->nlist_dirty = true;
    return LG_OK;
}

int lg_coulomb_wolf(lg_system* s, double cutoff, int restriction, double scale14) {
    if (!s) return fail(LG_E_INVALID, "null system");
    if (!(cutoff > 0.0)) return fail(LG_E_INVALID, "Got a negative cutoff in Wolf summation");
    LG_TRY(check_coulomb_cutoff(s, cutoff));
    if (restriction < 0 || restriction > 6) return fail(LG_E_INVALID, "unknown restriction");
    s->coul_kind = COUL_WOLF;
    s->coul_rc = cutoff;
    s->coul_alpha = kPi / cutoff;
    s->coul_kmax = 0;
    s->coul_restr = restriction;
    s->coul_s14 = scale14;
    s->params_dirty = s->kspace_dirty = s->nlist_dirty = true;
    return LG_OK;
}

int lg_coulomb_clear(lg_system* s) {
    if (!s) return fail(LG_E_INVALID, "null system");
    s->coul_kind = COUL_NONE;
    s->params_dirty = s->kspace_dirty = s->nlist_dirty = true;
    return LG_OK;
}

int lg_forces(lg_system* s, double* out, int accumulate) {
    if (!s || !out) return fail(LG_E_INVALID, "null argument");
    LG_HIP(hipSetDevice(s->ctx->device));
    LG_TRY(compute(s, false));
    return pull_vec3(s->force, s->n, out, S(s), accumulate != 0);
}

}  // extern "C"

namespace {
// EnergyEvaluator terms and their sum (energy.rs:47-158, compute.rs:114-127)
// from the gathered device sums.
int energy_terms_of(lg_system* s, const Energies& E, lg_energy_terms* out) {
    double e_tail, w_tail;
    tail_terms(s, &e_tail, &w_tail);
    double self = coulomb_self(s);
    lg_energy_terms t{};
    t.pairs = E.pairs;
    t.pairs_tail = e_tail;
    t.bonds = E.bonds;
    t.angles = E.angles;
    t.dihedrals = E.dihedrals;
    t.coulomb_real = E.coul_real;
    t.coulomb_self = self;
    t.coulomb_kspace = E.kspace;
    t.coulomb = s->coul_kind == COUL_NONE ? 0.0 : (E.coul_real + self + E.kspace);
    t.global = 0.0;
    double total = t.pairs;
    total += t.pairs_tail;
    total += t.bonds;
    total += t.angles;
    total += t.dihedrals;
    total += t.coulomb;
    total += t.global;
    t.total = total;
    *out = t;
    if (!std::isfinite(total)) return fail(LG_E_NONFINITE, "Potential energy is infinite!");
    return LG_OK;
}

// The 18 sums of gather_energies, added on the device in the same order
// (block order, pair blocks then intra-molecular blocks): one thread per sum.
__global__ void k_energy_gather(const double* __restrict__ pp, int nbp, const double* __restrict__ pm, int nbm,
                                const double* __restrict__ pk, int nbk, double* __restrict__ out) {
    const int q = threadIdx.x;
    if (q >= 18) return;
    double v = 0.0;
    if (q < 2 || q >= 6) {  // pairs, coulomb real space, w[0..5]: pair then intra blocks
        const int ip = q < 2 ? q : q - 6 + 2;
        const int im = q < 2 ? q : q - 6 + 5;
        if (q < 12) {
            for (int b = 0; b < nbp; b++) v += pp[b * 8 + ip];
            for (int b = 0; b < nbm; b++) v += pm[b * 11 + im];
        } else {  // wk[0..5]: k-space blocks
            for (int b = 0; b < nbk; b++) v += pk[b * 7 + 1 + (q - 12)];
        }
    } else if (q < 5) {  // bonds, angles, dihedrals: intra blocks
        for (int b = 0; b < nbm; b++) v += pm[b * 11 + q];
    } else {  // q == 5: k-space energy
        for (int b = 0; b < nbk; b++) v += pk[b * 7];
    }
    out[q] = v;
}
}  // namespace

extern "C" {

int lg_potential_energy(lg_system* s, lg_energy_terms* out) {
    if (!s || !out) return fail(LG_E_INVALID, "null argument");
    LG_HIP(hipSetDevice(s->ctx->device));
    LG_TRY(compute(s, true));
    Energies E;
    LG_TRY(gather_energies(s, E));
    return energy_terms_of(s, E, out);
}

int lg_potential_energy_noalloc(lg_system* s, lg_energy_terms* out) {
    if (!s || !out) return fail(LG_E_INVALID, "null argument");
    if (!s->ctx->scratch) return fail(LG_E_STATE, "the context has no scratch address");
    if (s->topo_dirty || s->meta_dirty || s->params_dirty || s->kspace_dirty || s->nlist_dirty || !s->epart_pairs.p)
        return fail(LG_E_STATE, "the system changed since its last evaluation: call lg_potential_energy (or "
                                "lg_forces) once first, which allocates what the evaluation needs");
    if (s->periodic) {  // the checks of prepare(), which allocates nothing
        double half = std::min(s->L[0], std::min(s->L[1], s->L[2])) / 2.0;
        for (auto& [k, p] : s->pairs)
            if (p.rc > half)
                return fail(LG_E_INVALID, "The cutoff for this pair potential is bigger than half the cell length");
        if (s->coul_kind != COUL_NONE && s->coul_rc > half)
            return fail(LG_E_INVALID, "The cutoff for the coulomb potential is bigger than half the cell length");
    }
    const int nbp = pair_blocks_of(s->n, s->tpa);
    const int nbm = s->nterms > 0 ? blocks_for(s->nterms, 256) : 0;
    const int nbk = (s->coul_kind == COUL_EWALD && s->nk > 0) ? blocks_for(s->nk, kRhoReduceK) : 0;
    const size_t need = size_t(nbp) * 8 + size_t(nbm) * 11 + size_t(nbk) * 7 + 18;
    if (need > kNoallocScratchDoubles)
        return fail(LG_E_INVALID, "the energy partial sums of this system (" + std::to_string(need) +
                                      " doubles) exceed the static scratch of " +
                                      std::to_string(kNoallocScratchDoubles) + " doubles");
    LG_HIP(hipSetDevice(s->ctx->device));
    double* base = s->ctx->scratch;
    s->ep_pairs_at = base;
    s->ep_mol_at = base + size_t(nbp) * 8;
    s->ep_k_at = s->ep_mol_at + size_t(nbm) * 11;
    double* sums = s->ep_k_at + size_t(nbk) * 7;
    const uint32_t mask = s->prof_mask;
    s->prof_mask = 0;  // no event creation in this call
    int rc = launch_nlist(s, true);
    if (rc == LG_OK) rc = launch_forces(s, true);
    s->prof_mask = mask;
    s->ep_pairs_at = s->ep_mol_at = s->ep_k_at = nullptr;
    if (rc != LG_OK) return rc;
    hipStream_t st = S(s);
    k_energy_gather<<<1, 64, 0, st>>>(base, nbp, base + size_t(nbp) * 8, nbm, base + size_t(nbp) * 8 + size_t(nbm) * 11,
                                      nbk, sums);
    LG_HIP(hipGetLastError());
    double h[18];
    int fl = 0;
    LG_HIP(hipMemcpyAsync(h, sums, sizeof(h), hipMemcpyDeviceToHost, st));
    LG_HIP(hipMemcpyAsync(&fl, s->flags.p + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    LG_HIP(hipStreamSynchronize(st));
    if (fl > 0)  // growing the neighbour list would allocate: leave that to lg_potential_energy
        return fail(LG_E_STATE, "the neighbour list is full: call lg_potential_energy once to grow it");
    Energies E;
    E.pairs = h[0];
    E.coul_real = h[1];
    E.bonds = h[2];
    E.angles = h[3];
    E.dihedrals = h[4];
    E.kspace = h[5] / kFourPiEps0;
    for (int q = 0; q < 6; q++) {
        E.w[q] = h[6 + q];
        E.wk[q] = h[12 + q] / kFourPiEps0;
    }
    return energy_terms_of(s, E, out);
}

int lg_virial(lg_system* s, double* w9) {
    if (!s || !w9) return fail(LG_E_INVALID, "null argument");
    if (!s->periodic) return fail(LG_E_INVALID, "Can not compute virial for infinite cell");
    LG_HIP(hipSetDevice(s->ctx->device));
    LG_TRY(compute(s, true));
    Energies E;
    LG_TRY(gather_energies(s, E));
    double e_tail, w_tail;
    tail_terms(s, &e_tail, &w_tail);
    double w[6];
    for (int q = 0; q < 6; q++) w[q] = E.w[q] + E.wk[q];
    w[0] += w_tail;
    w[1] += w_tail;
    w[2] += w_tail;
    double m[9] = {w[0], w[3], w[4], w[3], w[1], w[5], w[4], w[5], w[2]};
    std::memcpy(w9, m, sizeof(m));
    return LG_OK;
}

// MolecularVirial (compute.rs:278-364): pairs between molecules weighted by
// (r_ab . r_ij) / r_ab^2, the tails as the atomic virial, no bond, angle or
// dihedral term, and for Ewald the real-space pairs between molecules plus
// the atomic k-space virial minus sum_i F_i (x) (r_i - com) (ewald.rs:736-753).
int lg_molecular_virial(lg_system* s, double* w9) {
    if (!s || !w9) return fail(LG_E_INVALID, "null argument");
    if (!s->periodic) return fail(LG_E_INVALID, "Can not compute virial for infinite cell");
    LG_HIP(hipSetDevice(s->ctx->device));
    LG_TRY(compute(s, true));
    Energies E;
    LG_TRY(gather_energies(s, E));
    hipStream_t st = S(s);
    const int n = int(s->n);
    LG_TRY(s->mol_com.ensure(size_t(s->nmol)));
    k_mol_com<<<blocks_for(s->nmol, 256), 256, 0, st>>>(s->pos.p, s->vel.p, s->d_mol_start.p, s->nmol,
                                                        s->mol_com.p);
    const int nbp = blocks_for(n, 256);
    const bool kspace = s->coul_kind == COUL_EWALD && s->nk > 0;
    LG_TRY(s->molvir_part.ensure(size_t(nbp) * (6 + (kspace ? 9 : 0))));
    MolVirialArgs a{};
    a.pos = s->pos.p;
    a.meta = s->meta.p;
    a.sorted = s->sorted.p;
    a.nbr = s->nbr.p;
    a.nnbr = s->nnbr.p;
    a.cap = s->cap;
    a.n = n;
    a.nkinds = int(s->nkinds);
    a.table = s->table.p;
    a.tabs = s->ptabs.p;
    a.coul = s->coul;
    a.box = s->box;
    a.com = s->mol_com.p;
    a.part = s->molvir_part.p;
    k_mol_pairs<<<nbp, 256, 0, st>>>(a);
    if (kspace)
        k_mol_kspace<<<nbp, 256, 0, st>>>(s->pos.p, s->meta.p, s->fk.p, s->ksplit, n, s->mol_com.p,
                                          s->molvir_part.p + size_t(nbp) * 6);
    LG_HIP(hipGetLastError());
    std::vector<double> h(size_t(nbp) * (6 + (kspace ? 9 : 0)));
    LG_HIP(hipMemcpyAsync(h.data(), s->molvir_part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    LG_HIP(hipStreamSynchronize(st));
    double w6[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nbp; b++)
        for (int q = 0; q < 6; q++) w6[q] += h[size_t(b) * 6 + q];
    double e_tail, w_tail;
    tail_terms(s, &e_tail, &w_tail);
    double m[9] = {w6[0], w6[3], w6[4], w6[3], w6[1], w6[5], w6[4], w6[5], w6[2]};
    m[0] += w_tail;
    m[4] += w_tail;
    m[8] += w_tail;
    if (kspace) {
        double corr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const double* hk = h.data() + size_t(nbp) * 6;
        for (int b = 0; b < nbp; b++)
            for (int q = 0; q < 9; q++) corr[q] += hk[size_t(b) * 9 + q];
        const double wk[9] = {E.wk[0], E.wk[3], E.wk[4], E.wk[3], E.wk[1], E.wk[5], E.wk[4], E.wk[5], E.wk[2]};
        for (int q = 0; q < 9; q++) m[q] += wk[q] - corr[q];
    }
    std::memcpy(w9, m, sizeof(m));
    return LG_OK;
}

int lg_kinetic_energy(lg_system* s, double* out) {
    if (!s || !out) return fail(LG_E_INVALID, "null argument");
    std::vector<double> v(3 * s->n);
    LG_HIP(hipSetDevice(s->ctx->device));
    LG_TRY(pull_vec3(s->vel, s->n, v.data(), S(s)));
    double e = 0.0;  // compute.rs:137-142, sequential
    for (size_t i = 0; i < s->n; i++)
        e += 0.5 * s->mass[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
    *out = e;
    if (!std::isfinite(e)) return fail(LG_E_NONFINITE, "Kinetic energy is infinite!");
    return LG_OK;
}

int lg_temperature(lg_system* s, double* out) {
    double ke;
    LG_TRY(lg_kinetic_energy(s, &ke));
    *out = 2.0 * ke / (double(3 * s->n) * kBoltzmann);
    return LG_OK;
}

int lg_pairs_within(lg_system* s, double cutoff, uint32_t* pairs, uint64_t capacity, uint64_t* count) {
    if (!s || !count || (capacity && !pairs)) return fail(LG_E_INVALID, "null argument");
    LG_HIP(hipSetDevice(s->ctx->device));
    LG_TRY(prepare(s));
    if (!(cutoff > 0.0) || cutoff > s->rlist) return fail(LG_E_INVALID, "cutoff above the neighbour-list range");
    LG_TRY(compute(s, false));
    DevBuf<unsigned long long> cnt;
    DevBuf<uint2> out;
    LG_TRY(cnt.ensure(1));
    LG_TRY(out.ensure(std::max<uint64_t>(capacity, 1)));
    LG_HIP(hipMemsetAsync(cnt.p, 0, sizeof(unsigned long long), S(s)));
    int n = int(s->n);
    k_pairs_within<<<blocks_for(n, 256), 256, 0, S(s)>>>(s->pos.p, s->sorted.p, s->nbr.p, s->nnbr.p, n, s->cap,
                                                         s->box, cutoff, out.p, capacity, cnt.p);
    LG_HIP(hipGetLastError());
    unsigned long long c = 0;
    LG_HIP(hipMemcpyAsync(&c, cnt.p, sizeof(c), hipMemcpyDeviceToHost, S(s)));
    LG_HIP(hipStreamSynchronize(S(s)));
    uint64_t w = std::min<uint64_t>(c, capacity);
    if (w) LG_HIP(hipMemcpy(pairs, out.p, w * sizeof(uint2), hipMemcpyDeviceToHost));
    *count = c;
    cnt.release();
    out.release();
    return LG_OK;
}

int lg_ewald_kvector_count(lg_system* s, int* out) {
    if (!s || !out) return fail(LG_E_INVALID, "null argument");
    LG_HIP(hipSetDevice(s->ctx->device));
    LG_TRY(prepare(s));
    *out = s->nk;
    return LG_OK;
}

int lg_md_setup(lg_system* s, int integrator, double timestep, double bT, double btau) {
    if (!s) return fail(LG_E_INVALID, "null system");
    if (integrator != LG_INTEGRATOR_VELOCITY_VERLET && integrator != LG_INTEGRATOR_LEAP_FROG)
        return fail(LG_E_UNSUPPORTED, "unknown integrator");
    if (!(timestep > 0.0)) return fail(LG_E_INVALID, "timestep must be positive");
    // thermostats.rs:103-104: temperature >= 0 (0 included) and tau >= 1;
    // a negative or NaN temperature means no thermostat
    const bool thermo = bT >= 0.0;
    if (thermo && !std::isfinite(bT)) return fail(LG_E_INVALID, "The temperature must be finite in thermostats.");
    if (thermo && !(btau >= 1.0))
        return fail(LG_E_INVALID, "The timestep must be larger than 1 in berendsen thermostat.");
    s->integrator = integrator;
    s->dt = timestep;
    s->thermostat = thermo;
    s->berendsen_T = thermo ? bT : 0.0;
    s->berendsen_tau = thermo ? btau : 1.0;
    LG_HIP(hipSetDevice(s->ctx->device));
    // Integrator::setup: accelerations start at zero (integrators.rs:40-42)
    LG_HIP(hipMemsetAsync(s->acc.p, 0, s->n * sizeof(double4), S(s)));
    LG_HIP(hipStreamSynchronize(S(s)));
    s->md_ready = true;
    return LG_OK;
}

// The kernels of `todo` MD steps and, with a thermostat, the Berendsen
// scale at the end: launched eagerly, or captured into a HIP graph by
// lg_md_run.  No allocation, copy, memset or synchronisation in between.
static int md_chunk(lg_system* s, uint64_t todo) {
    hipStream_t st = S(s);
    const int n = int(s->n);
    const bool thermo = s->thermostat;
    const bool kspace = s->coul_kind == COUL_EWALD && s->nk > 0;
    double hs = 0.5 * s->skin;
    int nb = blocks_for(n, 64);  // second half: one wave per block, one KE partial each
    // the first half-step also writes the phase table of the new positions
    // (64 atoms per block, ld atoms with the padding)
    PhaseOut ph{};
    if (kspace) {
        ph.mode = s->precision == LG_PRECISION_MIXED ? 2 : 1;
        ph.kmax = s->coul_kmax;
        ph.ld = s->ld;
        ph.box = s->box;
        ph.tab = s->phase_tab.p;
        ph.q = s->phase_q.p;
        ph.tab32 = s->phase_tab32.p;
        ph.q32 = s->phase_q32.p;
    }
    Thermo th{nullptr, nb, double(3 * s->n), s->berendsen_T, s->berendsen_tau};
    int nbf = blocks_for(std::max(n, kspace ? s->ld : 0), 64);
    const int tbf = ph.mode == 2 ? 512 : 256;  // waves per first-half block: see phase_roles
    const CellTag tags = cell_tags(s);
    const bool lf = s->integrator != LG_INTEGRATOR_VELOCITY_VERLET;
    const int ks = kspace ? s->ksplit : 0;
    const int* sl = s->nterms > 0 ? s->slot_start.p : nullptr;
    // without a thermostat the second half of a step and the first half of
    // the next share a launch (k_md_fused); the chunk's first and last
    // halves run alone
    bool first_pending = true;
    for (uint64_t k = 0; k < todo; k++) {
        // integrator launches timed, when profiled, by their own kernel
        // timestamps (launch_timed), like the force launches
        if (first_pending) {
            // the Berendsen scale of the previous step is applied by the first half
            th.ke_part = (thermo && k > 0) ? s->ke_part.p : nullptr;
            if (!lf)
                launch_timed(s, LG_KERNEL_INTEGRATE, k_vv_first, nbf, tbf, 0, s->pos.p, s->vel.p, s->acc.p, n, s->dt,
                             th, s->ref.p, hs * hs, s->flags.p, ph, tags);
            else
                launch_timed(s, LG_KERNEL_INTEGRATE, k_lf_first, nbf, tbf, 0, s->pos.p, s->vel.p, s->acc.p, n, s->dt,
                             th, s->ref.p, hs * hs, s->flags.p, ph, tags);
        }
        LG_TRY(launch_nlist(s, false));
        LG_TRY(launch_forces(s, false, false, true));
        const bool fuse = !thermo && k + 1 < todo;
        if (fuse && !lf)
            launch_timed(s, LG_KERNEL_INTEGRATE, k_md_fused<false>, nbf, tbf, 0, s->force.p, s->pos.p, s->contrib.p,
                         sl, s->fk.p, ks, s->vel.p, s->acc.p, n, s->dt, s->ref.p, hs * hs, s->flags.p, ph, tags);
        else if (fuse)
            launch_timed(s, LG_KERNEL_INTEGRATE, k_md_fused<true>, nbf, tbf, 0, s->force.p, s->pos.p, s->contrib.p,
                         sl, s->fk.p, ks, s->vel.p, s->acc.p, n, s->dt, s->ref.p, hs * hs, s->flags.p, ph, tags);
        else {
            double* kp = thermo ? s->ke_part.p : nullptr;
            if (!lf)
                launch_timed(s, LG_KERNEL_INTEGRATE, k_md_second<false>, nb, 64, 0, s->force.p, s->pos.p,
                             s->contrib.p, sl, s->fk.p, ks, s->vel.p, s->acc.p, n, s->dt, kp, s->flags.p);
            else
                launch_timed(s, LG_KERNEL_INTEGRATE, k_md_second<true>, nb, 64, 0, s->force.p, s->pos.p,
                             s->contrib.p, sl, s->fk.p, ks, s->vel.p, s->acc.p, n, s->dt, kp, s->flags.p);
        }
        first_pending = !fuse;
    }
    if (thermo) {
        th.ke_part = s->ke_part.p;
        k_berendsen<<<1, 256, 0, st>>>(th, s->lambda.p);
        k_scale_vel<<<blocks_for(n, 256), 256, 0, st>>>(s->vel.p, n, s->lambda.p);
    }
    LG_HIP(hipGetLastError());
    return LG_OK;
}

static uint64_t bits_of(double x) {
    uint64_t u;
    std::memcpy(&u, &x, sizeof(u));
    return u;
}

// Steps in chunks; a chunk whose neighbour list overflowed is replayed from a
// snapshot with a larger list.
int lg_md_run(lg_system* s, uint64_t nsteps) {
    if (!s) return fail(LG_E_INVALID, "null system");
    if (!s->md_ready) return fail(LG_E_STATE, "lg_md_setup must be called before lg_md_run");
    LG_HIP(hipSetDevice(s->ctx->device));
    LG_TRY(prepare(s));
    s->mc_ready = s->mc_staged = s->mc_cells_valid = false;
    hipStream_t st = S(s);
    const uint64_t chunk = 200;
    LG_TRY(s->snap_pos.ensure(s->n));
    LG_TRY(s->snap_vel.ensure(s->n));
    LG_TRY(s->snap_acc.ensure(s->n));
    uint64_t done = 0;
    int retries = 0;
    LG_HIP(hipMemsetAsync(s->flags.p + 2, 0, sizeof(int), st));
    while (done < nsteps) {
        uint64_t todo = std::min<uint64_t>(chunk, nsteps - done);
        LG_HIP(hipMemcpyAsync(s->snap_pos.p, s->pos.p, s->n * sizeof(double4), hipMemcpyDeviceToDevice, st));
        LG_HIP(hipMemcpyAsync(s->snap_vel.p, s->vel.p, s->n * sizeof(double4), hipMemcpyDeviceToDevice, st));
        LG_HIP(hipMemcpyAsync(s->snap_acc.p, s->acc.p, s->n * sizeof(double4), hipMemcpyDeviceToDevice, st));
        const bool graph = s->tune.md_graph == 2 && s->prof_mask == 0;
        if (graph) {
            // captured once per (chunk length, configuration, buffer set,
            // integrator settings) and replayed
            const uint64_t key[8] = {todo, s->cfg_epoch, g_alloc_epoch.load(), uint64_t(s->integrator),
                                     uint64_t(s->precision), bits_of(s->dt),
                                     s->thermostat ? bits_of(s->berendsen_T) : ~uint64_t(0),
                                     bits_of(s->berendsen_tau)};
            if (!s->md_exec || !std::equal(key, key + 8, s->md_key)) {
                if (s->md_exec) (void)hipGraphExecDestroy(s->md_exec);
                s->md_exec = nullptr;
                hipGraph_t g = nullptr;
                LG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
                int rc = md_chunk(s, todo);
                hipError_t e = hipStreamEndCapture(st, &g);
                if (rc != LG_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (e != hipSuccess) return fail(LG_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
                e = hipGraphInstantiate(&s->md_exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess) {
                    s->md_exec = nullptr;
                    return fail(LG_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
                }
                std::copy(key, key + 8, s->md_key);
            }
            LG_HIP(hipGraphLaunch(s->md_exec, st));
        } else {
            LG_TRY(md_chunk(s, todo));
        }
        LG_HIP(hipGetLastError());
        bool ov = false, bad = false;
        LG_TRY(check_overflow(s, &ov, &bad));
        if (ov) {
            if (++retries > 4) return fail(LG_E_DEVICE, "neighbour list overflow could not be resolved");
            LG_HIP(hipMemcpyAsync(s->pos.p, s->snap_pos.p, s->n * sizeof(double4), hipMemcpyDeviceToDevice, st));
            LG_HIP(hipMemcpyAsync(s->vel.p, s->snap_vel.p, s->n * sizeof(double4), hipMemcpyDeviceToDevice, st));
            LG_HIP(hipMemcpyAsync(s->acc.p, s->snap_acc.p, s->n * sizeof(double4), hipMemcpyDeviceToDevice, st));
            LG_HIP(hipMemsetAsync(s->flags.p + 2, 0, sizeof(int), st));
            continue;
        }
        // a stronger form of the reference's periodic sanity_check
        // (simulations.rs:94-96, 129-155), which only warns: a run whose
        // positions or velocities are no longer finite fails, with the
        // state of the end of the chunk left on the device
        if (bad) return fail(LG_E_NONFINITE, "positions or velocities became non-finite during the MD run");
        done += todo;
    }
    LG_HIP(hipStreamSynchronize(st));
    return LG_OK;
}

int lg_profile_begin(lg_system* s, uint32_t mask) {
    if (!s) return fail(LG_E_INVALID, "null system");
    s->prof_mask = mask & ((1u << LG_KERNEL_CLASSES) - 1);
    s->prof_used = 0;
    s->prof_records.clear();
    return LG_OK;
}

int lg_profile_end(lg_system* s, lg_profile* out) {
    if (!s || !out) return fail(LG_E_INVALID, "null argument");
    LG_HIP(hipSetDevice(s->ctx->device));
    LG_HIP(hipStreamSynchronize(S(s)));
    lg_profile p{};
    for (auto& [cls, idx] : s->prof_records) {
        float ms = 0.0f;
        LG_HIP(hipEventElapsedTime(&ms, s->prof_events[idx], s->prof_events[idx + 1]));
        p.ms[cls] += double(ms);
        p.launches[cls] += 1;
    }
    s->prof_mask = 0;
    s->prof_used = 0;
    s->prof_records.clear();
    *out = p;
    return LG_OK;
}

int lg_cache_init(lg_system* s, double* energy) {
    if (!s) return fail(LG_E_INVALID, "null system");
    LG_HIP(hipSetDevice(s->ctx->device));
    LG_TRY(mc_init(s));
    if (energy) LG_TRY(mc_total(s, energy));
    return LG_OK;
}

int lg_cache_energy(lg_system* s, double* energy) {
    if (!s || !energy) return fail(LG_E_INVALID, "null argument");
    LG_HIP(hipSetDevice(s->ctx->device));
    LG_TRY(prepare(s));
    if (!s->mc_ready) LG_TRY(mc_init(s));
    return mc_total(s, energy);
}

int lg_move_cost(lg_system* s, uint32_t molid, const double* newpos, double* cost) {
    if (!s || !newpos || !cost) return fail(LG_E_INVALID, "null argument");
    LG_HIP(hipSetDevice(s->ctx->device));
    LG_TRY(prepare(s));
    if (!s->mc_ready) LG_TRY(mc_init(s));
    if (molid >= uint32_t(s->nmol)) return fail(LG_E_INVALID, "molecule id out of range");
    int m0 = s->mol_start[molid], m = s->mol_start[molid + 1] - m0;
    if (m > kMcMaxMol) return fail(LG_E_UNSUPPORTED, "molecule too large for a device Monte Carlo move");
    std::vector<double4> np(m);
    for (int a = 0; a < m; a++) {
        for (int d = 0; d < 3; d++)
            if (!std::isfinite(newpos[3 * a + d])) return fail(LG_E_NONFINITE, "non-finite position");
        np[a] = make_double4(newpos[3 * a], newpos[3 * a + 1], newpos[3 * a + 2], s->charge[m0 + a]);
    }
    hipStream_t st = S(s);
    LG_HIP(hipMemcpyAsync(s->mc_newpos.p, np.data(), m * sizeof(double4), hipMemcpyHostToDevice, st));
    McChain c = mc_chain(s, false);
    k_mc_cost<<<1, kMcBlock, mc_phase_bytes(s), st>>>(c, int(molid), s->mc_newpos.p);
    LG_HIP(hipGetLastError());
    double sums[4];
    LG_HIP(hipMemcpyAsync(sums, s->mc_out.p + 2, sizeof(sums), hipMemcpyDeviceToHost, st));
    LG_HIP(hipStreamSynchronize(st));
    double dp, dc;
    mc_terms(sums, dp, dc);
    s->mc_staged = true;
    s->mc_mol = molid;
    s->mc_dp = dp;
    s->mc_dc = dc;
    *cost = dp + dc;
    return LG_OK;
}

int lg_move_commit(lg_system* s, int wrap) {
    if (!s) return fail(LG_E_INVALID, "null system");
    if (!s->mc_staged) return fail(LG_E_STATE, "no staged move to commit");
    LG_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = S(s);
    McChain c = mc_chain(s, false);
    c.wrap = wrap != 0;
    k_mc_commit<<<1, kMcBlock, 0, st>>>(c, int(s->mc_mol), s->mc_newpos.p, s->mc_dp, s->mc_dc);
    LG_HIP(hipGetLastError());
    s->mc_staged = false;
    s->mc_cells_valid = false;
    LG_TRY(flag_rebuild(s));
    LG_HIP(hipStreamSynchronize(st));
    return LG_OK;
}

int lg_move_discard(lg_system* s) {
    if (!s) return fail(LG_E_INVALID, "null system");
    s->mc_staged = false;
    return LG_OK;
}

int lg_mc_run(lg_system* s, uint64_t seed, uint64_t attempt0, uint64_t nattempts, double beta, double delta,
              uint64_t* accepted) {
    if (!s) return fail(LG_E_INVALID, "null system");
    return lg_mc_run_batch(&s, 1, &seed, &attempt0, &nattempts, &beta, &delta, accepted);
}

int lg_mc_run_batch(lg_system* const* systems, size_t count, const uint64_t* seeds, const uint64_t* attempt0,
                    const uint64_t* nattempts, const double* beta, const double* delta, uint64_t* accepted) {
    if (!systems || !seeds || !attempt0 || !nattempts || !beta || !delta) return fail(LG_E_INVALID, "null argument");
    if (count == 0) return LG_OK;
    if (count > (size_t(1) << 20)) return fail(LG_E_INVALID, "too many chains in one batch");
    lg_context* ctx = systems[0] ? systems[0]->ctx : nullptr;
    if (!ctx) return fail(LG_E_INVALID, "null system");
    for (size_t b = 0; b < count; b++) {
        if (!systems[b]) return fail(LG_E_INVALID, "null system");
        if (systems[b]->ctx != ctx) return fail(LG_E_INVALID, "all chains of a batch must share one context");
        if (!(beta[b] >= 0.0) || !std::isfinite(beta[b]) || !(delta[b] >= 0.0) || !std::isfinite(delta[b]))
            return fail(LG_E_INVALID, "invalid temperature or amplitude");
    }
    {
        // duplicates by sorting (a pairwise check cost ~0.5 ms of host time
        // per launch of 1024 chains)
        std::vector<lg_system*> seen(systems, systems + count);
        std::sort(seen.begin(), seen.end());
        if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
            return fail(LG_E_INVALID, "a system appears twice in one batch");
    }
    LG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    std::vector<uint64_t> next(count), end(count), acc(count, 0);
    size_t shmem = 0;
    for (size_t b = 0; b < count; b++) {
        lg_system* s = systems[b];
        LG_TRY(prepare(s));
        if (!s->mc_ready) LG_TRY(mc_init(s));
        if (s->max_mol > kMcMaxMol) return fail(LG_E_UNSUPPORTED, "molecule too large for a device Monte Carlo move");
        if (!s->mc_cells_valid) LG_TRY(mc_prepare_cells(s, false));
        shmem = std::max(shmem, mc_phase_bytes(s));
        next[b] = attempt0[b];
        end[b] = attempt0[b] + nattempts[b];
    }
    if (shmem + sizeof(McShared) > ctx->max_shared) return fail(LG_E_UNSUPPORTED, "kmax too large for the MC LDS tables");
    LG_TRY(ctx->mc_desc.ensure(count));
    LG_TRY(ctx->mc_cnt.ensure(count * 2));
    std::vector<size_t> active(count);
    std::iota(active.begin(), active.end(), size_t(0));
    std::vector<McChain> desc;
    std::vector<unsigned long long> cnt;
    for (int round = 0; !active.empty(); round++) {
        if (round > 8) return fail(LG_E_DEVICE, "Monte Carlo cell list overflow could not be resolved");
        desc.clear();
        // chains with a Coulomb potential first: they run the kCoul kernel
        std::stable_partition(active.begin(), active.end(),
                              [&](size_t b) { return systems[b]->coul_kind != COUL_NONE; });
        size_t ncoul = 0;
        for (size_t b : active) ncoul += systems[b]->coul_kind != COUL_NONE;
        for (size_t b : active) {
            lg_system* s = systems[b];
            McChain c = mc_chain(s, true);
            c.key = make_uint2(uint32_t(seeds[b]), uint32_t(seeds[b] >> 32));
            c.attempt0 = next[b];
            c.nattempts = end[b] - next[b];
            c.beta = beta[b];
            c.d = delta[b] / std::sqrt(3.0);  // translate.rs:43
            c.wrap = 1;
            c.rebin = s->mc_cells_valid ? 0 : 1;
            // the cell list assumes a move stays within the neighbouring cells
            if (c.ncell > 0 && !(c.d < std::min(c.csize[0], std::min(c.csize[1], c.csize[2])))) {
                c.ncell = 0;
                s->mc_cells_valid = false;
            }
            size_t k = desc.size();
            c.counters = ctx->mc_cnt.p + 2 * k;
            desc.push_back(c);
        }
        size_t na = desc.size();
        LG_HIP(hipMemcpyAsync(ctx->mc_desc.p, desc.data(), na * sizeof(McChain), hipMemcpyHostToDevice, st));
        LG_HIP(hipMemsetAsync(ctx->mc_cnt.p, 0, na * 2 * sizeof(unsigned long long), st));
        if (ncoul > 0) k_mc_chains<true><<<unsigned(ncoul), kMcBlock, shmem, st>>>(ctx->mc_desc.p);
        if (na > ncoul) k_mc_chains<false><<<unsigned(na - ncoul), kMcBlock, shmem, st>>>(ctx->mc_desc.p + ncoul);
        LG_HIP(hipGetLastError());
        cnt.resize(na * 2);
        LG_HIP(hipMemcpyAsync(cnt.data(), ctx->mc_cnt.p, na * 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              st));
        LG_HIP(hipStreamSynchronize(st));
        std::vector<size_t> again;
        for (size_t k = 0; k < na; k++) {
            size_t b = active[k];
            lg_system* s = systems[b];
            acc[b] += cnt[2 * k];
            next[b] = cnt[2 * k + 1];
            s->mc_cells_valid = desc[k].ncell > 0;
            if (next[b] < end[b]) {
                // a cell would overflow: give every cell room for all atoms
                LG_TRY(mc_prepare_cells(s, true));
                again.push_back(b);
            }
            LG_TRY(flag_rebuild(s));
        }
        active.swap(again);
    }
    LG_HIP(hipStreamSynchronize(st));
    if (accepted)
        for (size_t b = 0; b < count; b++) accepted[b] = acc[b];
    return LG_OK;
}

}  // extern "C"
