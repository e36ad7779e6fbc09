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
// largest cutoff, >= 3 cells per axis) so that a move visits the 27 cells
// around the old and the new position instead of all atoms; smaller boxes use
// the all-atoms loop.  Ewald k-space: rho(k) is cached per chain and updated
// by the staged delta on accept, as ewald.rs:947-955.
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
constexpr int kMcMaxMol = 32;  // atoms per molecule that a move may displace

struct McChain {
    double4* pos;           // x, y, z, charge (updated on accept)
    const double4* vel;     // .w = mass (centre of mass for the wrap)
    const int2* meta;       // kind, molecule
    const int* mol_start;   // nmol + 1
    int n, nmol;
    Box box;
    const PairParam* table;
    int nkinds;
    CoulParam coul;
    // Ewald k-space (nk == 0: none)
    int kmax, nk;
    const int4* kidx;
    const double* efac;     // energy factor per k
    double2* rho;           // cached rho(k)
    double2* drho;          // staged delta rho(k) of the last cost evaluation
    // cell list (ncell == 0: all-atoms loop)
    int nc[3];
    double csize[3];
    int ncell, ccap;
    int* cell_atoms;        // ncell x ccap
    int* cell_count;
    int* atom_cell;
    int* atom_slot;
    // run
    uint2 key;
    unsigned long long attempt0, nattempts;
    double beta, d;         // 1 / kT, per-axis half width delta / sqrt(3)
    int wrap;
    int rebin;              // rebuild the cell list before the first attempt
    // results: out[0] delta pairs, out[1] delta coulomb, out[2..7] the six
    // partial sums of the last cost; counters[0] accepted, counters[1] the
    // attempt at which the run stopped (attempt0 + nattempts when complete)
    double* out;
    unsigned long long* counters;
};

// Pair energy of an inter-molecular pair at distance r: the pair potential
// (cut at r >= rc, shifted) and the real-space Coulomb term (cut at r > rc),
// as k_pairs; restrictions reduce to the flags for "different molecules".
__device__ __forceinline__ void mc_pair_energy(const PairParam* tab, int nkinds, const CoulParam& c, int ki, int kj,
                                               double qiqj, double r, double& ep, double& ec) {
    double inv_r = 1.0 / r;
    const PairParam& pp = tab[ki * nkinds + kj];
    if (pp.pot > POT_NULL && !pp.inter_excl && r < pp.rc) {
        double e, f;
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

__device__ __forceinline__ int mc_cell_of(const McChain& c, double4 p) {
    int cx = cell_coord(p.x, c.box.L[0], c.nc[0], c.csize[0], c.box.periodic);
    int cy = cell_coord(p.y, c.box.L[1], c.nc[1], c.csize[1], c.box.periodic);
    int cz = cell_coord(p.z, c.box.L[2], c.nc[2], c.csize[2], c.box.periodic);
    return (cz * c.nc[1] + cy) * c.nc[0] + cx;
}

__device__ __forceinline__ int mc_neighbour_cell(const McChain& c, int cell, int q) {
    int cx = cell % c.nc[0], cy = (cell / c.nc[0]) % c.nc[1], cz = cell / (c.nc[0] * c.nc[1]);
    int dx = q % 3 - 1, dy = (q / 3) % 3 - 1, dz = q / 9 - 1;
    int x = cx + dx, y = cy + dy, z = cz + dz;
    x = x < 0 ? x + c.nc[0] : (x >= c.nc[0] ? x - c.nc[0] : x);
    y = y < 0 ? y + c.nc[1] : (y >= c.nc[1] ? y - c.nc[1] : y);
    z = z < 0 ? z + c.nc[2] : (z >= c.nc[2] ? z - c.nc[2] : z);
    return (z * c.nc[1] + y) * c.nc[0] + x;
}

struct McShared {
    PairParam tab[kMaxKinds * kMaxKinds];
    double4 newp[kMcMaxMol];
    double red[kMcBlock / 64][6];
    double sums[6];
    int seg_cell[54];
    int seg_start[55];
    int flag;
};

// Block-cooperative energy change of moving molecule `mol` (atoms m0..m0+m-1)
// to sh.newp: sums[0..5] = pairs old/new, coulomb real old/new, k-space
// old/new (the last two before the 1 / (4 pi eps0) factor).  Stages
// drho[k] for every k.  `ph` is dynamic LDS for the molecule's phase rows.
__device__ void mc_delta(const McChain& c, McShared& sh, double2* ph, int mol, int m0, int m) {
    int tid = threadIdx.x;
    double eo_p = 0, en_p = 0, eo_c = 0, en_c = 0;
    for (int a = 0; a < m; a++) {
        double4 po = c.pos[m0 + a];
        double4 pn = sh.newp[a];
        int ki = c.meta[m0 + a].x;
        if (c.ncell == 0) {
            for (int j = tid; j < c.n; j += kMcBlock) {
                if (j >= m0 && j < m0 + m) continue;
                double4 pj = c.pos[j];
                int kj = c.meta[j].x;
                double qq = po.w * pj.w;
                mc_pair_energy(sh.tab, c.nkinds, c.coul, ki, kj, qq, mc_pair_r(c.box, po, pj), eo_p, eo_c);
                mc_pair_energy(sh.tab, c.nkinds, c.coul, ki, kj, qq, mc_pair_r(c.box, pn, pj), en_p, en_c);
            }
        } else {
            // the 27 cells around the old position, then around the new one
            if (tid < 54) {
                int home = mc_cell_of(c, tid < 27 ? po : pn);
                int cell = mc_neighbour_cell(c, home, tid % 27);
                sh.seg_cell[tid] = cell;
                sh.seg_start[tid + 1] = c.cell_count[cell];
            }
            __syncthreads();
            if (tid == 0) {
                sh.seg_start[0] = 0;
                for (int q = 0; q < 54; q++) sh.seg_start[q + 1] += sh.seg_start[q];
            }
            __syncthreads();
            int total = sh.seg_start[54];
            for (int idx = tid; idx < total; idx += kMcBlock) {
                int lo = 0, hi = 54;  // seg_start[lo] <= idx < seg_start[hi]
                while (hi - lo > 1) {
                    int mid = (lo + hi) >> 1;
                    if (sh.seg_start[mid] <= idx)
                        lo = mid;
                    else
                        hi = mid;
                }
                int j = c.cell_atoms[(size_t)sh.seg_cell[lo] * c.ccap + (idx - sh.seg_start[lo])];
                if (j >= m0 && j < m0 + m) continue;
                double4 pj = c.pos[j];
                int kj = c.meta[j].x;
                double qq = po.w * pj.w;
                if (lo < 27)
                    mc_pair_energy(sh.tab, c.nkinds, c.coul, ki, kj, qq, mc_pair_r(c.box, po, pj), eo_p, eo_c);
                else
                    mc_pair_energy(sh.tab, c.nkinds, c.coul, ki, kj, qq, mc_pair_r(c.box, pn, pj), en_p, en_c);
            }
            __syncthreads();
        }
    }
    double ko = 0, kn = 0;
    if (c.nk > 0) {
        // phase rows of the old (w = 0) and new (w = 1) positions, by the
        // recursion of ewald.rs:646-661 (as k_phase)
        int R = phase_rows(c.kmax);
        for (int task = tid; task < m * 6; task += kMcBlock) {
            int a = task / 6, w = (task / 3) % 2, axis = task % 3;
            double4 p = w == 0 ? c.pos[m0 + a] : sh.newp[a];
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
                double q = c.pos[m0 + a].w;
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
    double v[6] = {eo_p, en_p, eo_c, en_c, ko, kn};
    int lane = tid & 63, wid = tid >> 6;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        double s = wave_sum(v[q]);
        if (lane == 0) sh.red[wid][q] = s;
    }
    __syncthreads();
    if (tid < 6) {
        double s = 0;
        for (int w = 0; w < kMcBlock / 64; w++) s += sh.red[w][tid];
        sh.sums[tid] = s;
    }
    __syncthreads();
}

// Energy change terms from the six sums: pairs, and coulomb = real + k-space
// (ewald.rs:812-836 divides each k-space sum by 4 pi eps0 before the
// difference).
__host__ __device__ __forceinline__ void mc_terms(const double* s, double& d_pairs, double& d_coul) {
    d_pairs = s[1] - s[0];
    d_coul = (s[3] - s[2]) + (s[5] / kFourPiEps0 - s[4] / kFourPiEps0);
}

#pragma clang fp contract(off)
// Final positions of an accepted move: the new positions, then the centre of
// mass wrapped into the cell (translate.rs:115-127, molecules.rs:287-296),
// with the reference's operation order.  Runs on one thread.
__device__ void mc_final_positions(const McChain& c, McShared& sh, int m0, int m) {
    if (!c.wrap || !c.box.periodic) return;
    double total = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
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
    for (int a = 0; a < m; a++) {
        sh.newp[a].x += dx;
        sh.newp[a].y += dy;
        sh.newp[a].z += dz;
    }
}
#pragma clang fp contract(on)

// True when the cells of the final positions can take the molecule's atoms.
__device__ bool mc_cells_fit(const McChain& c, McShared& sh, int m0, int m) {
    if (c.ncell == 0) return true;
    for (int a = 0; a < m; a++) {
        int cell = mc_cell_of(c, sh.newp[a]);
        if (cell != c.atom_cell[m0 + a] && c.cell_count[cell] + m > c.ccap) return false;
    }
    return true;
}

// Moves atom i from its cell to `cell` (swap with the last entry).
__device__ void mc_rebin(const McChain& c, int i, int cell) {
    int old = c.atom_cell[i];
    if (old == cell) return;
    int slot = c.atom_slot[i];
    int last = c.cell_count[old] - 1;
    int moved = c.cell_atoms[(size_t)old * c.ccap + last];
    c.cell_atoms[(size_t)old * c.ccap + slot] = moved;
    c.atom_slot[moved] = slot;
    c.cell_count[old] = last;
    int s = c.cell_count[cell];
    c.cell_atoms[(size_t)cell * c.ccap + s] = i;
    c.cell_count[cell] = s + 1;
    c.atom_cell[i] = cell;
    c.atom_slot[i] = s;
}

__device__ __forceinline__ void mc_load_table(const McChain& c, McShared& sh) {
    for (int x = threadIdx.x; x < c.nkinds * c.nkinds; x += kMcBlock) sh.tab[x] = c.table[x];
}

// Cost of moving molecule `mol` to newpos[0..m) (one block).  out[2..7] get
// the six sums; drho is staged for k_mc_commit.
__global__ __launch_bounds__(kMcBlock) void k_mc_cost(McChain c, int mol, const double4* newpos) {
    extern __shared__ double2 ph[];
    __shared__ McShared sh;
    mc_load_table(c, sh);
    int m0 = c.mol_start[mol], m = c.mol_start[mol + 1] - m0;
    if (threadIdx.x < m) sh.newp[threadIdx.x] = newpos[threadIdx.x];
    __syncthreads();
    mc_delta(c, sh, ph, mol, m0, m);
    if (threadIdx.x < 6) c.out[2 + threadIdx.x] = sh.sums[threadIdx.x];
}

// Applies the staged move: positions (with the centre-of-mass wrap when
// c.wrap) and rho += drho (one block).
__global__ __launch_bounds__(kMcBlock) void k_mc_commit(McChain c, int mol, const double4* newpos) {
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
    for (int k = threadIdx.x; k < c.nk; k += kMcBlock) {
        double2 r = c.rho[k], d = c.drho[k];
        c.rho[k] = make_double2(r.x + d.x, r.y + d.y);
    }
}

// Bins every atom of the chain, in index order (one block, serial insert so
// that the order inside a cell is deterministic).  counters[1] is set to 0
// when a cell would exceed its capacity.
__device__ bool mc_bin_all(const McChain& c) {
    if (c.ncell == 0) return true;
    __shared__ int ok;
    for (int x = threadIdx.x; x < c.ncell; x += kMcBlock) c.cell_count[x] = 0;
    for (int i = threadIdx.x; i < c.n; i += kMcBlock) c.atom_cell[i] = mc_cell_of(c, c.pos[i]);
    __syncthreads();
    if (threadIdx.x == 0) {
        ok = 1;
        for (int i = 0; i < c.n; i++) {
            int cell = c.atom_cell[i];
            int s = c.cell_count[cell];
            if (s >= c.ccap) {
                ok = 0;
                break;
            }
            c.cell_atoms[(size_t)cell * c.ccap + s] = i;
            c.cell_count[cell] = s + 1;
            c.atom_slot[i] = s;
        }
    }
    __syncthreads();
    return ok != 0;
}

#pragma clang fp contract(off)
// The Metropolis chain: block b runs chains[b] for its attempts
// (monte_carlo.rs:289-330 with Translate, translate.rs:85-127), with the
// displacement arithmetic unfused as in the reference.  c.rebin rebuilds the
// chain's cell list first.  A chain stops early, with counters[1] = the
// attempt not yet made, when a cell list would overflow.
__global__ __launch_bounds__(kMcBlock) void k_mc_chains(const McChain* __restrict__ chains) {
    extern __shared__ double2 ph[];
    __shared__ McShared sh;
    const McChain c = chains[blockIdx.x];
    mc_load_table(c, sh);
    if (threadIdx.x == 0) sh.flag = 0;
    __syncthreads();
    bool binned = c.rebin ? mc_bin_all(c) : true;
    double d_pairs_total = 0.0, d_coul_total = 0.0;
    unsigned long long accepted = 0;
    unsigned long long t = c.attempt0, end = c.attempt0 + c.nattempts;
    if (!binned) end = t;
    for (; t < end; t++) {
        uint4 r0 = philox4x32_10(make_uint4(uint32_t(t), uint32_t(t >> 32), 0u, 0u), c.key);
        uint4 r1 = philox4x32_10(make_uint4(uint32_t(t), uint32_t(t >> 32), 1u, 0u), c.key);
        uint4 r2 = philox4x32_10(make_uint4(uint32_t(t), uint32_t(t >> 32), 2u, 0u), c.key);
        int mol = int(u01(r0.x, r0.y) * double(c.nmol));
        if (mol >= c.nmol) mol = c.nmol - 1;
        double ux = u01(r0.z, r0.w), uy = u01(r1.x, r1.y), uz = u01(r1.z, r1.w);
        double uacc = u01(r2.x, r2.y);
        double dx = -c.d + 2.0 * c.d * ux, dy = -c.d + 2.0 * c.d * uy, dz = -c.d + 2.0 * c.d * uz;
        int m0 = c.mol_start[mol], m = c.mol_start[mol + 1] - m0;
        if (threadIdx.x < m) {
            double4 p = c.pos[m0 + threadIdx.x];
            sh.newp[threadIdx.x] = make_double4(p.x + dx, p.y + dy, p.z + dz, p.w);
        }
        __syncthreads();
        mc_delta(c, sh, ph, mol, m0, m);
        if (threadIdx.x == 0) {
            double dp, dc;
            mc_terms(sh.sums, dp, dc);
            double cost = c.beta * (dp + dc);
            int accept = (cost <= 0.0 || uacc < exp(-cost)) ? 1 : 0;
            if (accept) {
                mc_final_positions(c, sh, m0, m);
                if (!mc_cells_fit(c, sh, m0, m)) {
                    accept = 2;  // stop before this attempt
                } else {
                    d_pairs_total += dp;
                    d_coul_total += dc;
                    accepted++;
                    for (int a = 0; a < m; a++) {
                        double4 p = sh.newp[a];
                        c.pos[m0 + a] = make_double4(p.x, p.y, p.z, p.w);
                        if (c.ncell) mc_rebin(c, m0 + a, mc_cell_of(c, p));
                    }
                }
            }
            sh.flag = accept;
        }
        __syncthreads();
        int acc = sh.flag;
        if (acc == 2) break;
        if (acc == 1)
            for (int k = threadIdx.x; k < c.nk; k += kMcBlock) {
                double2 r = c.rho[k], d = c.drho[k];
                c.rho[k] = make_double2(r.x + d.x, r.y + d.y);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        c.out[0] += d_pairs_total;
        c.out[1] += d_coul_total;
        c.counters[0] += accepted;
        c.counters[1] = t;
    }
}
#pragma clang fp contract(on)

__global__ void k_set_flag(int* p, int v) { *p = v; }

}  // namespace lg
