// lg_mc.hpp -- Monte Carlo translation moves on gfx950 (FP64).
//
// Reference: EnergyCache (lumol-core/src/sys/cache.rs:75-220),
// SharedEwald/Ewald MC cost (energy/global/ewald.rs:572-617, :756-839,
// :933-955), Wolf MC cost (energy/global/wolf.rs:121-165), the Metropolis
// loop (lumol-sim/src/mc/monte_carlo.rs:289-330) and Translate
// (lumol-sim/src/mc/moves/translate.rs:41-148).
//
// Design.  One workgroup owns one Markov chain and runs its attempts back to
// back on the device.  The reference caches an N x N pair-energy matrix; here
// the old pair energies are recomputed from the current positions instead.
// They are bit-identical to the cached ones (the minimum image of -d is the
// negated minimum image of d, and the norm squares it), so the cost differs
// from the reference's only through the order of the summation.  The chain's
// atoms are binned in a cell list of fixed-capacity cells (cell edge >= the
// largest cutoff, >= 4 cells per axis) holding copies of the positions, so
// that a move reads the cells around its old and new position instead of all
// atoms; smaller boxes use the all-atoms loop.  Ewald k-space: rho(k) is
// cached per chain and updated by the staged delta on accept, as
// ewald.rs:947-955.
//
// A chain is a serial dependency (attempt t + 1 sees attempt t's outcome), so
// the kernel is built for latency: every thread evaluates the Metropolis test
// from the same reduced sums (no broadcast, no barrier after a rejection),
// the molecule of attempt t + 1 is drawn and its positions loaded while
// attempt t runs, and the per-wave partial sums are double-buffered so one
// barrier per attempt separates the waves.
//
// Random numbers: Philox4x32-10 keyed by the chain seed; attempt t uses the
// counters (t, 0), (t, 1), (t, 2) for the molecule choice, the displacement
// and the acceptance draw, so a chain is reproducible from (seed, t) alone.
#pragma once

#include "lg_kernels.hpp"

namespace lg {

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ------------------------------
__host__ __device__ __forceinline__ uint32_t mulhi32(uint32_t a, uint32_t b) {
    return uint32_t((uint64_t(a) * uint64_t(b)) >> 32);
}

__host__ __device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
    for (int r = 0; r < 10; r++) {
        uint32_t hi0 = mulhi32(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        uint32_t hi1 = mulhi32(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

// 53-bit uniform double in [0, 1) from two 32-bit words.
__host__ __device__ __forceinline__ double u01(uint32_t a, uint32_t b) {
    uint64_t x = (uint64_t(a) << 21) ^ uint64_t(b >> 11);
    x &= ((1ull << 53) - 1);
    return double(x) * (1.0 / 9007199254740992.0);
}

// ---- chain descriptor ----------------------------------------------------------------
constexpr int kMcBlock = 256;
constexpr int kMcWaves = kMcBlock / 64;
constexpr int kMcMaxMol = 32;     // atoms per molecule that a move may displace
constexpr int kMcMaxCells = 4096; // cells of a chain (their counts live in LDS)

struct McChain {
    double4* pos;           // x, y, z, charge (updated on accept)
    const double4* vel;     // .w = mass (centre of mass for the wrap)
    const int2* meta;       // kind, molecule
    const int* mol_start;   // nmol + 1
    int n, nmol;
    Box box;
    const PairParam* table;
    const double* tabs;     // TableComputation tables of the pair potentials
    int nkinds;
    CoulParam coul;
    // Ewald k-space (nk == 0: none)
    int kmax, nk;
    const int4* kidx;
    const double* efac;     // energy factor per k
    double2* rho;           // cached rho(k)
    double2* drho;          // staged delta rho(k) of the last cost evaluation
    // cell list of the chain kernel (ncell == 0: all-atoms loop): each cell
    // holds up to ccap entries, a copy of the atom's position and charge and
    // its (index, kind)
    int nc[3];
    double csize[3];
    double inv_csize[3];    // 1 / csize
    int ncell, ccap;
    double4* cell_pos;      // ncell x ccap
    int2* cell_ids;         // ncell x ccap: atom, kind
    int* cell_count;        // ncell (kept in LDS while a chain runs)
    int* atom_cell;         // n (binning scratch)
    // run
    double r2_filter;       // (largest cutoff)^2 * (1 + 1e-6): candidates beyond skip the exact path
    uint2 key;
    unsigned long long attempt0, nattempts;
    double beta, d;         // 1 / kT, per-axis half width delta / sqrt(3)
    int wrap;
    int rebin;              // rebuild the cell list before the first attempt
    // out[0], out[1]: the cached pair and coulomb energies, updated by every
    // accepted move in attempt order (so a run gives the same bits however it
    // is split into launches); out[2..5]: the four sums of the last cost.
    // counters[0] accepted, counters[1] the attempt at which the run stopped
    // (attempt0 + nattempts when complete)
    double* out;
    unsigned long long* counters;
};

// Pair energy of an inter-molecular pair at distance r: the pair potential
// (cut at r >= rc, shifted) and the real-space Coulomb term (cut at r > rc),
// as k_pairs; restrictions reduce to the flags for "different molecules".
// `tab` holds the fast fields; an entry marked POT_GENERAL is evaluated from
// the full table `full` (and the TableComputation tables `tabs`).
__device__ __forceinline__ void mc_pair_energy(const PairFast* tab, const PairParam* full, const double* tabs,
                                               int nkinds, const CoulParam& c, int ki, int kj, double qiqj, double r,
                                               double& ep, double& ec) {
    double inv_r = 1.0 / r;
    const PairFast& pp = tab[ki * nkinds + kj];
    if (pp.pot > POT_NULL && !pp.inter_excl && r < pp.rc) {
        double e, f;
        if (pp.pot == POT_GENERAL)
            pot_eval_any(full[ki * nkinds + kj], tabs, r, inv_r, e, f);
        else
            pot_eval(pp.pot, pp.p0, pp.p1, r, inv_r, e, f);
        ep += e - pp.shift;
    }
    if (c.kind != COUL_NONE && qiqj != 0.0 && r <= c.rc) {
        double e, f;
        if (c.kind == COUL_EWALD) {
            ewald_real(c.alpha, qiqj, r, inv_r, c.inter_excl != 0, e, f);
            ec += e;
        } else if (!c.inter_excl) {
            wolf_pair(c, qiqj, r, inv_r, e, f);
            ec += e;
        }
    }
}

__device__ __forceinline__ double mc_pair_r(const Box& b, double4 p, double4 q) {
    double dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    min_image(b, dx, dy, dz);
    return sqrt(norm2_ordered(dx, dy, dz));
}

// Squared distance with a plain nearest-integer image.  It can pick another
// image than the reference's only for a component within rounding of L/2,
// where both images have the same length up to rounding; with the 1e-6
// relative margin of r2_filter, `r2_near < r2_filter` keeps every pair that
// the reference's distance puts within the cutoff.
__device__ __forceinline__ double mc_r2_near(const Box& b, double4 p, double4 q) {
    double dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
    if (b.periodic) {
        dx -= rint(dx * b.inv[0]) * b.L[0];
        dy -= rint(dy * b.inv[1]) * b.L[1];
        dz -= rint(dz * b.inv[2]) * b.L[2];
    }
    return dx * dx + dy * dy + dz * dz;
}

// Energy change of one candidate j: e(new) - e(old), each evaluated with the
// reference's arithmetic when the candidate passes the distance filter.
template <bool kCoul>
__device__ __forceinline__ void mc_candidate(const McChain& c, const PairFast* tab, const CoulParam& cp, int ki,
                                             int kj, double4 po, double4 pn, double4 pj, double& d_p,
                                             double& d_c) {
    double qq = kCoul ? po.w * pj.w : 0.0;
    double eo_p = 0, eo_c = 0, en_p = 0, en_c = 0;
    if (mc_r2_near(c.box, po, pj) < c.r2_filter)
        mc_pair_energy(tab, c.table, c.tabs, c.nkinds, cp, ki, kj, qq, mc_pair_r(c.box, po, pj), eo_p, eo_c);
    if (mc_r2_near(c.box, pn, pj) < c.r2_filter)
        mc_pair_energy(tab, c.table, c.tabs, c.nkinds, cp, ki, kj, qq, mc_pair_r(c.box, pn, pj), en_p, en_c);
    d_p += en_p - eo_p;
    if (kCoul) d_c += en_c - eo_c;
}

// q = v / e, r = v % e for e in {3, 4} (constant divisors).
__device__ __forceinline__ void mc_divmod34(int v, int e, int& q, int& r) {
    q = e == 4 ? (v >> 2) : v / 3;
    r = v - q * e;
}

// Cell coordinate of x along one axis of the chain's cell list: wrapped
// into [0, L) and scaled by the inverse cell edge.  Only self-consistency
// matters (binning and lookups use this one function); the cells are
// larger than the cutoff by a relative 1e-12, far above the rounding.
__host__ __device__ __forceinline__ int mc_axis_cell(double x, double L, double invL, double inv_cs, int nc) {
    double w = x - floor(x * invL) * L;
    return clamp_cell(floor(w * inv_cs), nc);
}

__device__ __forceinline__ int3 mc_cell3(const McChain& c, double4 p) {
    return make_int3(mc_axis_cell(p.x, c.box.L[0], c.box.inv[0], c.inv_csize[0], c.nc[0]),
                     mc_axis_cell(p.y, c.box.L[1], c.box.inv[1], c.inv_csize[1], c.nc[1]),
                     mc_axis_cell(p.z, c.box.L[2], c.box.inv[2], c.inv_csize[2], c.nc[2]));
}

__device__ __forceinline__ int mc_cell_of(const McChain& c, double4 p) {
    int3 q = mc_cell3(c, p);
    return (q.z * c.nc[1] + q.y) * c.nc[0] + q.x;
}

struct McShared {
    PairFast tab[kMaxKinds * kMaxKinds];
    double4 newp[kMcMaxMol];       // trial positions (then final, on accept)
    double4 oldp[kMcMaxMol];
    double4 committed[kMcMaxMol];  // positions of the last accepted move
    double red[2][kMcWaves][4];    // per-wave partial sums, two buffers
    int self_slot[kMcMaxMol];
    int flag;
};

__device__ __forceinline__ void mc_load_table(const McChain& c, McShared& sh) {
    for (int x = threadIdx.x; x < c.nkinds * c.nkinds; x += kMcBlock) sh.tab[x] = pair_fast(c.table[x]);
}

// ---- energy change of a move -----------------------------------------------------------
// Candidates over the cell list.  The cells visited are the union box of the
// 27 cells around the old and the new cell of the atom (3 or 4 cells per
// axis; the host enables the cell list only with >= 4 cells per axis and a
// displacement below one cell, so each cell is visited once); thread t takes
// cell t / G of the box and its slots t % G, t % G + G, ... (G = 256 / cells
// in the box).  Every candidate adds e_new - e_old.  The thread that meets
// the atom itself records its slot in sh.self_slot[a].
template <bool kCoul>
__device__ __forceinline__ void mc_cells_atom(const McChain& c, McShared& sh, const int* cnt, int a, int m0, int m,
                                              double4 po, double4 pn, double& d_p, double& d_c) {
    int tid = threadIdx.x;
    int3 ho = mc_cell3(c, po), hn = mc_cell3(c, pn);
    int hov[3] = {ho.x, ho.y, ho.z}, hnv[3] = {hn.x, hn.y, hn.z};
    int lo[3], ext[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int delta = hnv[d] - hov[d];
        if (delta > 1) delta -= c.nc[d];
        if (delta < -1) delta += c.nc[d];
        lo[d] = hov[d] + (delta < 0 ? delta : 0) - 1;
        ext[d] = 3 + (delta != 0 ? 1 : 0);
    }
    int nseg = ext[0] * ext[1] * ext[2];
    // G = 256 / nseg and seg = tid / G by multiply-shift (exact for tid < 256;
    // nseg is 27, 36, 48 or 64)
    int G = nseg == 27 ? 9 : (nseg == 36 ? 7 : (nseg == 48 ? 5 : 4));
    int M = nseg == 27 ? 7282 : (nseg == 36 ? 9363 : (nseg == 48 ? 13108 : 16384));
    int seg = (tid * M) >> 16, sub = tid - seg * G;
    if (seg >= nseg) return;
    int sx, sy, sz, rest;
    mc_divmod34(seg, ext[0], rest, sx);
    mc_divmod34(rest, ext[1], sz, sy);
    int x = lo[0] + sx, y = lo[1] + sy, z = lo[2] + sz;
    x = x < 0 ? x + c.nc[0] : (x >= c.nc[0] ? x - c.nc[0] : x);
    y = y < 0 ? y + c.nc[1] : (y >= c.nc[1] ? y - c.nc[1] : y);
    z = z < 0 ? z + c.nc[2] : (z >= c.nc[2] ? z - c.nc[2] : z);
    int cell = (z * c.nc[1] + y) * c.nc[0] + x;
    int count = cnt[cell];
    int ki = c.meta[m0 + a].x;
    int self = m0 + a;
    CoulParam none{};
    const CoulParam& cp = kCoul ? c.coul : none;
    const size_t base = (size_t)cell * c.ccap;
    auto visit = [&](int slot, int2 id, double4 pj) {
        if (id.x >= m0 && id.x < m0 + m) {
            if (id.x == self) sh.self_slot[a] = slot;
            return;
        }
        mc_candidate<kCoul>(c, sh.tab, cp, ki, id.y, po, pn, pj, d_p, d_c);
    };
    // two slots per pass, both loads issued before either is used
#pragma unroll 1
    for (int slot = sub; slot < count; slot += 2 * G) {
        int s1 = slot + G;
        bool two = s1 < count;
        int2 id0 = c.cell_ids[base + slot];
        double4 p0 = c.cell_pos[base + slot];
        int2 id1 = two ? c.cell_ids[base + s1] : id0;
        double4 p1 = two ? c.cell_pos[base + s1] : p0;
        visit(slot, id0, p0);
        if (two) visit(s1, id1, p1);
    }
}

// Candidates over all atoms (chains without a cell list, and the per-move
// entry points).
template <bool kCoul>
__device__ __forceinline__ void mc_all_atom(const McChain& c, McShared& sh, int a, int m0, int m, double4 po,
                                            double4 pn, double& d_p, double& d_c) {
    int ki = c.meta[m0 + a].x;
    CoulParam none{};
    const CoulParam& cp = kCoul ? c.coul : none;
#pragma unroll 1
    for (int j = threadIdx.x; j < c.n; j += kMcBlock) {
        if (j >= m0 && j < m0 + m) continue;
        mc_candidate<kCoul>(c, sh.tab, cp, ki, c.meta[j].x, po, pn, c.pos[j], d_p, d_c);
    }
}

// Ewald k-space change of moving the molecule from sh.oldp to sh.newp
// (ewald.rs:756-839): phase rows of both positions by the recursion of
// ewald.rs:646-661 (as k_phase) in the dynamic LDS `ph`, then per k the
// delta rho (staged in drho[k]) and factor * |rho|^2 before and after.
__device__ __forceinline__ void mc_kspace(const McChain& c, McShared& sh, double2* ph, int m, double& ko,
                                          double& kn) {
    int tid = threadIdx.x;
    int R = phase_rows(c.kmax);
    for (int task = tid; task < m * 6; task += kMcBlock) {
        int a = task / 6, w = (task / 3) % 2, axis = task % 3;
        double4 p = w == 0 ? sh.oldp[a] : sh.newp[a];
        double x = axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
        double phase = (2.0 * kPi * c.box.inv[axis]) * x;
        double sn, cs;
        sincos(phase, &sn, &cs);
        double2* row = ph + (size_t)(a * 2 + w) * R;
        double2 e1 = make_double2(cs, sn);
        row[phase_row(axis, 0, c.kmax)] = make_double2(1.0, 0.0);
        row[phase_row(axis, 1, c.kmax)] = e1;
        if (axis != 0) row[phase_row(axis, -1, c.kmax)] = make_double2(e1.x, -e1.y);
        double2 e = e1;
        for (int k = 2; k <= c.kmax; k++) {
            e = make_double2(e.x * e1.x - e.y * e1.y, e.x * e1.y + e.y * e1.x);
            row[phase_row(axis, k, c.kmax)] = e;
            if (axis != 0) row[phase_row(axis, -k, c.kmax)] = make_double2(e.x, -e.y);
        }
    }
    __syncthreads();
    for (int k = tid; k < c.nk; k += kMcBlock) {
        int4 ix = c.kidx[k];
        int rx = phase_row(0, ix.x, c.kmax), ry = phase_row(1, ix.y, c.kmax), rz = phase_row(2, ix.z, c.kmax);
        double dre = 0.0, dim = 0.0;
        for (int a = 0; a < m; a++) {
            const double2* ro = ph + (size_t)(a * 2) * R;
            const double2* rn = ro + R;
            double2 op = cmul(cmul(ro[rx], ro[ry]), ro[rz]);
            double2 np = cmul(cmul(rn[rx], rn[ry]), rn[rz]);
            double q = sh.oldp[a].w;
            dre += q * (np.x - op.x);
            dim += q * (np.y - op.y);
        }
        c.drho[k] = make_double2(dre, dim);
        double2 r = c.rho[k];
        double f = c.efac[k];
        ko += f * (r.x * r.x + r.y * r.y);
        double nre = r.x + dre, nim = r.y + dim;
        kn += f * (nre * nre + nim * nim);
    }
}

// Per-wave partials of the sums into red[wave][0..NV) (pairs delta,
// coulomb real delta, k-space old, k-space new), then a barrier.
template <int NV>
__device__ __forceinline__ void mc_reduce(double (*red)[4], const double* v) {
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double s = wave_sum(v[q]);
        if (lane == 0) red[wid][q] = s;
    }
    __syncthreads();
}

// The four sums, adding the wave partials in wave order (any thread; every
// thread gets the same bits).
template <int NV>
__device__ __forceinline__ void mc_sums(const double (*red)[4], double* s) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double t = 0.0;
        if (q < NV)
            for (int w = 0; w < kMcWaves; w++) t += red[w][q];
        s[q] = t;
    }
}

// Energy change terms from the sums: pairs, and coulomb = real + k-space
// (ewald.rs:812-836 divides each k-space sum by 4 pi eps0 before the
// difference).
__host__ __device__ __forceinline__ void mc_terms(const double* s, double& d_pairs, double& d_coul) {
    d_pairs = s[0];
    d_coul = s[1] + (s[3] / kFourPiEps0 - s[2] / kFourPiEps0);
}

#pragma clang fp contract(off)
// Final positions of an accepted move: the new positions, then the centre of
// mass wrapped into the cell (translate.rs:115-127, molecules.rs:287-296),
// with the reference's operation order.  Runs on one thread.
__device__ __forceinline__ void mc_final_positions(const McChain& c, McShared& sh, int m0, int m) {
    if (!c.wrap || !c.box.periodic) return;
    double total = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
#pragma unroll 1
    for (int a = 0; a < m; a++) {
        double mass = c.vel[m0 + a].w;
        double4 p = sh.newp[a];
        total += mass;
        cx += mass * p.x;
        cy += mass * p.y;
        cz += mass * p.z;
    }
    cx = cx / total;
    cy = cy / total;
    cz = cz / total;
    double wx = cx - floor(cx / c.box.L[0]) * c.box.L[0];
    double wy = cy - floor(cy / c.box.L[1]) * c.box.L[1];
    double wz = cz - floor(cz / c.box.L[2]) * c.box.L[2];
    double dx = wx - cx, dy = wy - cy, dz = wz - cz;
#pragma unroll 1
    for (int a = 0; a < m; a++) {
        sh.newp[a].x += dx;
        sh.newp[a].y += dy;
        sh.newp[a].z += dz;
    }
}
#pragma clang fp contract(on)

// ---- per-move entry points (lg_move_cost / lg_move_commit) -------------------------
// Cost of moving molecule `mol` to newpos[0..m) (one block, all-atoms loop).
// out[2..5] get the four sums; drho is staged for k_mc_commit.
__global__ __launch_bounds__(kMcBlock) void k_mc_cost(McChain c, int mol, const double4* newpos) {
    extern __shared__ double2 ph[];
    __shared__ McShared sh;
    mc_load_table(c, sh);
    int m0 = c.mol_start[mol], m = c.mol_start[mol + 1] - m0;
    if (threadIdx.x < m) {
        sh.newp[threadIdx.x] = newpos[threadIdx.x];
        sh.oldp[threadIdx.x] = c.pos[m0 + threadIdx.x];
    }
    __syncthreads();
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int a = 0; a < m; a++) mc_all_atom<true>(c, sh, a, m0, m, sh.oldp[a], sh.newp[a], v[0], v[1]);
    if (c.nk > 0) mc_kspace(c, sh, ph, m, v[2], v[3]);
    mc_reduce<4>(sh.red[0], v);
    if (threadIdx.x == 0) {
        double s[4];
        mc_sums<4>(sh.red[0], s);
        for (int q = 0; q < 4; q++) c.out[2 + q] = s[q];
    }
}

// Applies the staged move: positions (with the centre-of-mass wrap when
// c.wrap), rho += drho and the cached energies += (dp, dc) (one block).
__global__ __launch_bounds__(kMcBlock) void k_mc_commit(McChain c, int mol, const double4* newpos, double dp,
                                                       double dc) {
    __shared__ McShared sh;
    int m0 = c.mol_start[mol], m = c.mol_start[mol + 1] - m0;
    if (threadIdx.x < m) sh.newp[threadIdx.x] = newpos[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) mc_final_positions(c, sh, m0, m);
    __syncthreads();
    if (threadIdx.x < m) {
        double4 p = sh.newp[threadIdx.x];
        c.pos[m0 + threadIdx.x] = make_double4(p.x, p.y, p.z, c.pos[m0 + threadIdx.x].w);
    }
    if (threadIdx.x == 0) {
        c.out[0] += dp;
        c.out[1] += dc;
    }
    for (int k = threadIdx.x; k < c.nk; k += kMcBlock) {
        double2 r = c.rho[k], d = c.drho[k];
        c.rho[k] = make_double2(r.x + d.x, r.y + d.y);
    }
}

// ---- the chain kernel ------------------------------------------------------------------
// Bins every atom of the chain into its cell, in index order (serial insert by
// thread 0, so the order inside a cell is deterministic); counts go to `cnt`
// (LDS).  False when a cell would exceed its capacity.
__device__ bool mc_bin_all(const McChain& c, int* cnt) {
    __shared__ int ok;
    for (int x = threadIdx.x; x < c.ncell; x += kMcBlock) cnt[x] = 0;
    for (int i = threadIdx.x; i < c.n; i += kMcBlock) c.atom_cell[i] = mc_cell_of(c, c.pos[i]);
    __syncthreads();
    if (threadIdx.x == 0) {
        ok = 1;
        for (int i = 0; i < c.n; i++) {
            int cell = c.atom_cell[i];
            int s = cnt[cell];
            if (s >= c.ccap) {
                ok = 0;
                break;
            }
            c.cell_pos[(size_t)cell * c.ccap + s] = c.pos[i];
            c.cell_ids[(size_t)cell * c.ccap + s] = make_int2(i, c.meta[i].x);
            cnt[cell] = s + 1;
        }
    }
    __syncthreads();
    return ok != 0;
}

// Applies an accepted move to the cell list (thread 0): same cell -> the
// entry's position copy is updated in place; otherwise the entry is removed
// (the cell's last entry takes its slot) and appended to the new cell.
__device__ __forceinline__ void mc_commit_cells(const McChain& c, McShared& sh, int* cnt, int m0, int m) {
#pragma unroll 1
    for (int a = 0; a < m; a++) {
        double4 p = sh.newp[a];
        int co = mc_cell_of(c, sh.oldp[a]), cn = mc_cell_of(c, p);
        int slot = sh.self_slot[a];
        size_t e = (size_t)co * c.ccap + slot;
        if (co == cn) {
            c.cell_pos[e] = p;
            continue;
        }
        int last = cnt[co] - 1;
        int2 me = c.cell_ids[e];
        if (slot != last) {
            size_t el = (size_t)co * c.ccap + last;
            int2 mid = c.cell_ids[el];
            c.cell_ids[e] = mid;
            c.cell_pos[e] = c.cell_pos[el];
            if (mid.x >= m0 && mid.x < m0 + m) sh.self_slot[mid.x - m0] = slot;
        }
        cnt[co] = last;
        int s = cnt[cn];
        size_t en = (size_t)cn * c.ccap + s;
        c.cell_ids[en] = me;
        c.cell_pos[en] = p;
        cnt[cn] = s + 1;
    }
}

__device__ __forceinline__ uint4 mc_draw(const McChain& c, unsigned long long t, uint32_t w) {
    return philox4x32_10(make_uint4(uint32_t(t), uint32_t(t >> 32), w, 0u), c.key);
}

__device__ __forceinline__ int mc_mol_of(const McChain& c, uint4 r0) {
    int mol = int(u01(r0.x, r0.y) * double(c.nmol));
    return mol >= c.nmol ? c.nmol - 1 : mol;
}

#pragma clang fp contract(off)
// The Metropolis chain: block b runs chains[b] for its attempts
// (monte_carlo.rs:289-330 with Translate, translate.rs:85-127), with the
// displacement arithmetic unfused as in the reference.  c.rebin rebuilds the
// chain's cell list first.  A chain stops early, with counters[1] = the
// attempt not yet made, when a cell would overflow.  kCoul: the chain has a
// Coulomb potential (else only the pair sums are formed).
template <bool kCoul>
__global__ __launch_bounds__(kMcBlock) void k_mc_chains(const McChain* __restrict__ chains) {
    extern __shared__ double2 ph[];
    __shared__ McShared sh;
    __shared__ int cnt[kMcMaxCells];
    const McChain& c = chains[blockIdx.x];
    int tid = threadIdx.x;
    mc_load_table(c, sh);
    bool cells = c.ncell > 0;
    bool ewald = kCoul && c.nk > 0;
    bool binned = true;
    if (cells) {
        if (c.rebin) {
            binned = mc_bin_all(c, cnt);
        } else {
            for (int x = tid; x < c.ncell; x += kMcBlock) cnt[x] = c.cell_count[x];
        }
    }
    __syncthreads();
    double e_pairs = c.out[0], e_coul = c.out[1];  // running cached energies (used by thread 0)
    unsigned long long accepted = 0;
    unsigned long long t = c.attempt0, end = c.attempt0 + c.nattempts;
    if (!binned) end = t;
    // the molecule of attempt t + 1 is drawn while attempt t runs and its
    // positions are loaded early: thread a < m holds atom a (all threads
    // hold atom 0, so single atoms need no LDS copy)
    uint4 r0 = mc_draw(c, t, 0u);
    int mol = t < end ? mc_mol_of(c, r0) : 0;
    int m0 = c.mol_start[mol], m = c.mol_start[mol + 1] - m0;
    double4 pre = c.pos[m0 + (tid < m ? tid : 0)];
    for (; t < end; t++) {
        uint4 r1 = mc_draw(c, t, 1u);
        uint4 r2 = mc_draw(c, t, 2u);
        double ux = u01(r0.z, r0.w), uy = u01(r1.x, r1.y), uz = u01(r1.z, r1.w);
        double uacc = u01(r2.x, r2.y);
        double dx = -c.d + 2.0 * c.d * ux, dy = -c.d + 2.0 * c.d * uy, dz = -c.d + 2.0 * c.d * uz;
        double4 po = pre;
        double4 pn = make_double4(po.x + dx, po.y + dy, po.z + dz, po.w);
        bool lds = m > 1 || ewald;
        if (tid < m) {
            sh.oldp[tid] = po;
            sh.newp[tid] = pn;
        }
        uint4 r0_next = mc_draw(c, t + 1, 0u);
        int mol_next = t + 1 < end ? mc_mol_of(c, r0_next) : mol;
        int n0 = c.mol_start[mol_next], nm = c.mol_start[mol_next + 1] - n0;
        pre = c.pos[n0 + (tid < nm ? tid : 0)];
        if (lds) __syncthreads();
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        if (!lds) {
            if (cells)
                mc_cells_atom<kCoul>(c, sh, cnt, 0, m0, 1, po, pn, v[0], v[1]);
            else
                mc_all_atom<kCoul>(c, sh, 0, m0, 1, po, pn, v[0], v[1]);
        } else {
            for (int a = 0; a < m; a++) {
                if (cells)
                    mc_cells_atom<kCoul>(c, sh, cnt, a, m0, m, sh.oldp[a], sh.newp[a], v[0], v[1]);
                else
                    mc_all_atom<kCoul>(c, sh, a, m0, m, sh.oldp[a], sh.newp[a], v[0], v[1]);
            }
            if (ewald) mc_kspace(c, sh, ph, m, v[2], v[3]);
        }
        int buf = int(t & 1);
        double s[4];
        if (kCoul) {
            mc_reduce<4>(sh.red[buf], v);
            mc_sums<4>(sh.red[buf], s);
        } else {
            mc_reduce<1>(sh.red[buf], v);
            mc_sums<1>(sh.red[buf], s);
        }
        double dp, dc;
        mc_terms(s, dp, dc);
        double cost = c.beta * (dp + dc);
        if (cost <= 0.0 || uacc < exp(-cost)) {
            // every thread takes this branch: the sums and draws are the same
            if (tid == 0) {
                int ok = 1;
                mc_final_positions(c, sh, m0, m);
                if (cells)
#pragma unroll 1
                    for (int a = 0; a < m; a++) {
                        int cell = mc_cell_of(c, sh.newp[a]);
                        if (cell != mc_cell_of(c, sh.oldp[a]) && cnt[cell] + m > c.ccap) ok = 0;
                    }
                if (ok) {
                    e_pairs += dp;
                    e_coul += dc;
                    accepted++;
#pragma unroll 1
                    for (int a = 0; a < m; a++) {
                        c.pos[m0 + a] = sh.newp[a];
                        sh.committed[a] = sh.newp[a];
                    }
                    if (cells) mc_commit_cells(c, sh, cnt, m0, m);
                }
                sh.flag = ok;
            }
            __syncthreads();
            if (!sh.flag) break;  // a cell would overflow: stop before attempt t
            if (ewald)
                for (int k = tid; k < c.nk; k += kMcBlock) {
                    double2 r = c.rho[k], dr = c.drho[k];
                    c.rho[k] = make_double2(r.x + dr.x, r.y + dr.y);
                }
            // the early copy of a molecule moved just now is stale
            if (mol_next == mol) pre = sh.committed[tid < nm ? tid : 0];
        }
        mol = mol_next;
        m0 = n0;
        m = nm;
        r0 = r0_next;
    }
    __syncthreads();
    if (cells && binned)
        for (int x = tid; x < c.ncell; x += kMcBlock) c.cell_count[x] = cnt[x];
    if (tid == 0) {
        c.out[0] = e_pairs;
        c.out[1] = e_coul;
        c.counters[0] += accepted;
        c.counters[1] = t;
    }
}
#pragma clang fp contract(on)

__global__ void k_set_flag(int* p, int v) { *p = v; }

}  // namespace lg
