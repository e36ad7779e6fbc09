// lg_kernels.hpp -- gfx950 kernels of the FP64 force-and-energy path.
//
// Data layout in HBM (per system, N atoms, original atom order):
//   pos   double4[N]  x, y, z, charge
//   vel   double4[N]  vx, vy, vz, mass
//   acc   double4[N]  ax, ay, az, -
//   meta  int2[N]     kind, molecule id
//   force double4[N]  fx, fy, fz, -
// Neighbour search: a cell list (integer binning + per-cell rank sort) feeds
// a Verlet list of inter-molecular candidates within rc + skin, stored
// atom-major (nbr[slot * cap + k]): the lanes sharing one atom read
// consecutive words of its row.
// The list is rebuilt on the device whenever an atom has moved more than
// skin/2 since the last build, which keeps every pair with r <= rc in the
// list; each kernel then applies the reference's own cutoff test on the
// FP64 distance, so the evaluated pair set is the one of the reference's
// brute-force `for j in i+1..n` loops (lumol-core/src/sys/compute.rs:40).
#pragma once

#include <type_traits>

#include "lg_device.hpp"

namespace lg {

// ---------------------------------------------------------------------------------
// Neighbour-list maintenance
// ---------------------------------------------------------------------------------
struct Grid {
    int n[3];           // cells per dimension
    double size[3];     // cell edge: >= rlist (reach 1), >= rlist / 2 (reach 2), or L (1 cell, reach 0)
    double inv_size[3]; // 1 / size
    int reach[3];       // neighbour cells searched on each side
    int ncell;
};

// A Verlet-list entry: the neighbour's atom index in bits 0-23 and its kind
// in bits 24-30, so that the pair kernels need no per-pair load of the kind.
__host__ __device__ __forceinline__ int nbr_entry(int j, int kind) { return j | (kind << 24); }
__host__ __device__ __forceinline__ int nbr_atom(int e) { return e & 0xffffff; }
__host__ __device__ __forceinline__ int nbr_kind(int e) { return e >> 24; }

// Rebuild protocol: a displacement test (here, or in the MD first half-step)
// sets *flag when an atom moved more than skin/2 since the last build; the
// list kernels that follow rebuild when it is set, and the pair kernel,
// which always runs after them and before the next test, clears it.  Stream
// order does all the signalling: no kernel has to detect the end of a
// rebuild (a grid-wide last-block handshake costs an L2 write-back per block
// on a multi-XCD device), and a step captured in a HIP graph needs no
// per-step arguments.
// Coordinate wrapped into [0, L) (periodic) as the cell binning sees it,
// and its cell index.  Multiplications by the inverses instead of
// divisions: the results move by a few ulps at most, far inside the 1e-12
// relative margin of the cell edges over rlist (and the 1e-4 margin of the
// FP32 candidate test), and the binning runs at FP64 multiply rate (the
// single-workgroup cell list of a small system bins every atom on one CU).
__host__ __device__ __forceinline__ double wrap_coord(double x, double L, double invL, int periodic) {
    return periodic ? x - floor(x * invL) * L : x;
}

// The cell index is clamped in floating point before the conversion, so that
// any input -- NaN, infinities, coordinates far outside the box -- gives an
// index in [0, nc) (NaN compares false and lands in cell 0): a simulation that
// blows up keeps every list kernel in bounds.
__host__ __device__ __forceinline__ int clamp_cell(double f, int nc) {
    return f >= 1.0 ? (f < (double)nc ? (int)f : nc - 1) : 0;
}

__host__ __device__ __forceinline__ int cell_coord(double x, double L, double invL, int nc, double inv_csize,
                                                   int periodic) {
    if (nc == 1) return 0;
    double w = wrap_coord(x, L, invL, periodic);
    return clamp_cell(floor(w * inv_csize), nc);
}

// Cell coordinates of every atom for the direct Verlet build, written by the
// kernels that see each new position (the displacement test, the first half
// of an MD step): cx | cy << 10 | cz << 20 (cells per axis <= 256).  Off
// when tag is null.
// What the MD half-steps write beside a new position: the cell tag of the
// direct Verlet build (tag, or null) and the position as the mixed pair
// kernel reads it (frac, or null; see frac_pos).
struct CellTag {
    int* tag;
    Grid g;
    Box b;
    float4* frac;
    double3 fo;  // origin of the FP32 Cartesian coordinates in an infinite cell
};
__host__ __device__ __forceinline__ int cell_tag(const CellTag& t, double4 p) {
    const int cx = cell_coord(p.x, t.b.L[0], t.b.inv[0], t.g.n[0], t.g.inv_size[0], t.b.periodic);
    const int cy = cell_coord(p.y, t.b.L[1], t.b.inv[1], t.g.n[1], t.g.inv_size[1], t.b.periodic);
    const int cz = cell_coord(p.z, t.b.L[2], t.b.inv[2], t.g.n[2], t.g.inv_size[2], t.b.periodic);
    return cx | (cy << 10) | (cz << 20);
}

// Position of an atom as the mixed pair kernel reads it, 16 bytes instead
// of 32 (the pair kernel's gathers move half the bytes).  In a periodic
// cell: the fractional coordinate x / L - floor(x / L), formed in FP64, as a
// 32-bit fixed-point fraction (units of L / 2^32, stored in the bits of the
// float lanes); the difference of two such words as a signed 32-bit integer
// is then the minimum image, exactly, at a resolution of L / 2^32, and the
// one rounding of the pair distance is its conversion to FP32, as when an
// FP64 difference is rounded.  In an infinite cell: FP32 Cartesian
// coordinates relative to an origin `fo` (the first atom's position when
// the positions were last set), formed in FP64, so that a cluster far from
// the coordinate origin keeps the precision of one near it.  The charge
// rides in w.
__device__ __forceinline__ unsigned fixed_frac(double x, double inv) {
    double t = x * inv;
    t = (t - floor(t)) * 4294967296.0;                   // [0, 2^32]
    return (unsigned)(unsigned long long)(long long)rint(t);  // 2^32 wraps to 0
}
__device__ __forceinline__ float4 frac_pos(const Box& b, double3 fo, double4 p) {
    if (!b.periodic) return make_float4((float)(p.x - fo.x), (float)(p.y - fo.y), (float)(p.z - fo.z), (float)p.w);
    return make_float4(__uint_as_float(fixed_frac(p.x, b.inv[0])), __uint_as_float(fixed_frac(p.y, b.inv[1])),
                       __uint_as_float(fixed_frac(p.z, b.inv[2])), (float)p.w);
}

__global__ void k_frac(const double4* __restrict__ pos, int n, Box b, double3 fo, float4* __restrict__ frac) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) frac[i] = frac_pos(b, fo, pos[i]);
}

__global__ void k_nlist_check(const double4* __restrict__ pos, const double4* __restrict__ ref, int n,
                              double half_skin2, int* __restrict__ flag, CellTag tg) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool moved = false;
    if (tg.tag && i < n) tg.tag[i] = cell_tag(tg, pos[i]);
    if (i < n) {
        double dx = pos[i].x - ref[i].x;
        double dy = pos[i].y - ref[i].y;
        double dz = pos[i].z - ref[i].z;
        moved = dx * dx + dy * dy + dz * dz > half_skin2;
    }
    if (__any(moved) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

__global__ void k_cell_count(const double4* __restrict__ pos, int n, Grid g, Box b, int* __restrict__ cellid,
                             int* __restrict__ slot, int* __restrict__ count, const int* __restrict__ flag) {
    if (*flag == 0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 p = pos[i];
    int cx = cell_coord(p.x, b.L[0], b.inv[0], g.n[0], g.inv_size[0], b.periodic);
    int cy = cell_coord(p.y, b.L[1], b.inv[1], g.n[1], g.inv_size[1], b.periodic);
    int cz = cell_coord(p.z, b.L[2], b.inv[2], g.n[2], g.inv_size[2], b.periodic);
    int c = (cz * g.n[1] + cy) * g.n[0] + cx;
    cellid[i] = c;
    slot[i] = atomicAdd(&count[c], 1);
}

// Exclusive scan of the per-cell counts (one block, 1024 threads), chunk by
// chunk of blockDim cells with a running carry; the counts of kScanBatch
// chunks are loaded together, so that the chunks wait for memory once per
// batch (at 50k cells, one wait per chunk cost ~300 us per rebuild).
constexpr int kScanBatch = 8;
__global__ void k_cell_scan(const int* __restrict__ count, int* __restrict__ start, int ncell, const int* __restrict__ flag) {
    if (*flag == 0) return;
    __shared__ int warp_sums[16];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int base0 = 0; base0 < ncell; base0 += kScanBatch * blockDim.x) {
        int vb[kScanBatch];
#pragma unroll
        for (int u = 0; u < kScanBatch; u++) {
            const int c = base0 + u * blockDim.x + threadIdx.x;
            vb[u] = c < ncell ? count[c] : 0;
        }
#pragma unroll
        for (int u = 0; u < kScanBatch; u++) {
            const int base = base0 + u * blockDim.x;
            if (base >= ncell) break;  // uniform
            const int c = base + threadIdx.x;
            const int v = vb[u];
            // inclusive scan within the wave
            int x = v;
            for (int off = 1; off < 64; off <<= 1) {
                int y = __shfl_up(x, off, 64);
                if (lane >= off) x += y;
            }
            if (lane == 63) warp_sums[wid] = x;
            __syncthreads();
            if (threadIdx.x == 0) {
                int s = 0;
                int nw = blockDim.x >> 6;
                for (int w = 0; w < nw; w++) {
                    int t = warp_sums[w];
                    warp_sums[w] = s;
                    s += t;
                }
            }
            __syncthreads();
            int excl = carry + warp_sums[wid] + x - v;
            if (c < ncell) start[c] = excl;
            __syncthreads();
            if (threadIdx.x == blockDim.x - 1) carry = excl + v;
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) start[ncell] = carry;
}

__global__ void k_cell_scatter(const int* __restrict__ cellid, const int* __restrict__ slot,
                               const int* __restrict__ start, int* __restrict__ tmp, int n, const int* __restrict__ flag) {
    if (*flag == 0) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tmp[start[cellid[i]] + slot[i]] = i;
}

// Deterministic order inside each cell: atoms by increasing index; an
// atom's rank is the number of smaller indices in its cell.  One wave per
// cell over a grid of at most kRankBlocks blocks that strides over the
// cells: a launch of one block per cell (50k cells at 1M atoms) cost ~70 us
// on every step, rebuild or not, in block dispatch alone.  Also writes the
// candidate record of the list build in sorted order: FP32 wrapped position
// (the binning's own wrap) and the molecule id.
constexpr int kRankBlocks = 4096;
__global__ __launch_bounds__(256) void k_cell_rank(const int* __restrict__ tmp, const int* __restrict__ start,
                                                   int* __restrict__ sorted, float4* __restrict__ cpos,
                                                   const double4* __restrict__ pos, const int2* __restrict__ meta,
                                                   Box b, int ncell, const int* __restrict__ flag) {
    if (*flag == 0) return;
    const int lane = threadIdx.x & 63;
    const int wid = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), nw = (int)((gridDim.x * blockDim.x) >> 6);
    for (int c = wid; c < ncell; c += nw) {
        const int s = start[c], e = start[c + 1];
        for (int a = s + lane; a < e; a += 64) {
            int v = tmp[a];
            int rank = 0;
            for (int b2 = s; b2 < e; b2++) rank += tmp[b2] < v;
            sorted[s + rank] = v;
            double4 p = pos[v];
            cpos[s + rank] = make_float4((float)wrap_coord(p.x, b.L[0], b.inv[0], b.periodic),
                                         (float)wrap_coord(p.y, b.L[1], b.inv[1], b.periodic),
                                         (float)wrap_coord(p.z, b.L[2], b.inv[2], b.periodic),
                                         __int_as_float(meta[v].y));
        }
    }
}

// Cell list of a small system in one workgroup of 1024 threads: the counts,
// offsets, scatter and per-cell rank order of the four-kernel chain above
// (k_cell_count .. k_cell_rank), with the counters, the slot table and the
// atoms' cells in LDS (dynamic: 8 ncell + 4 + 8 n bytes).  Same outputs:
// cellid[n], start[ncell + 1], sorted[n], cpos[n]; one launch instead of
// four, which is also what a step without a rebuild pays.  ncell < 65536.
__global__ __launch_bounds__(1024) void k_cell_small(const double4* __restrict__ pos, const int2* __restrict__ meta,
                                                     int n, Grid g, Box b, int* __restrict__ cellid,
                                                     int* __restrict__ start, int* __restrict__ sorted,
                                                     float4* __restrict__ cpos, const int* __restrict__ flag) {
    if (*flag == 0) return;
    extern __shared__ int smem_cells[];
    const int nc = g.ncell;
    int* cnt = smem_cells;                                         // [nc]: counts, then slot counters
    int* off = cnt + nc;                                           // [nc + 1]
    int* tmp = off + nc + 1;                                       // [n]: atom of each slot
    unsigned short* slot_cell = reinterpret_cast<unsigned short*>(tmp + n);  // [n]
    unsigned short* atom_cell = slot_cell + n;                               // [n]
    __shared__ int wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    for (int c = tid; c < nc; c += 1024) cnt[c] = 0;
    __syncthreads();
    // four atoms per thread per pass, their loads issued together
    for (int i0 = tid; i0 < n; i0 += 4 * 1024) {
        double4 pv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            int i = i0 + u * 1024;
            pv[u] = pos[i < n ? i : n - 1];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            int i = i0 + u * 1024;
            if (i < n) {
                int cx = cell_coord(pv[u].x, b.L[0], b.inv[0], g.n[0], g.inv_size[0], b.periodic);
                int cy = cell_coord(pv[u].y, b.L[1], b.inv[1], g.n[1], g.inv_size[1], b.periodic);
                int cz = cell_coord(pv[u].z, b.L[2], b.inv[2], g.n[2], g.inv_size[2], b.periodic);
                int c = (cz * g.n[1] + cy) * g.n[0] + cx;
                cellid[i] = c;
                atom_cell[i] = (unsigned short)c;
                atomicAdd(&cnt[c], 1);
            }
        }
    }
    __syncthreads();
    // exclusive scan of the counts: `per` consecutive cells per thread
    const int per = (nc + 1023) / 1024;
    const int c0 = tid * per, c1 = min(c0 + per, nc);
    int local = 0;
    for (int c = c0; c < c1; c++) local += cnt[c];
    int x = local;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (tid == 0) {
        int run = 0;
        for (int w = 0; w < 16; w++) {
            int t = wsum[w];
            wsum[w] = run;
            run += t;
        }
    }
    __syncthreads();
    int run = wsum[wid] + x - local;
    for (int c = c0; c < c1; c++) {
        int k = cnt[c];
        off[c] = run;
        start[c] = run;
        cnt[c] = run;
        run += k;
    }
    if (tid == 0) {
        off[nc] = n;
        start[nc] = n;
    }
    __syncthreads();
    // scatter in arbitrary order within a cell ...
    for (int i = tid; i < n; i += 1024) {
        int c = atom_cell[i];
        int sl = atomicAdd(&cnt[c], 1);
        tmp[sl] = i;
        slot_cell[sl] = (unsigned short)c;
    }
    __syncthreads();
    // ... made deterministic: atoms by increasing index within each cell
    for (int s0 = tid; s0 < n; s0 += 4 * 1024) {
        int vv[4], dst[4];
        double4 pv[4];
        int mv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            int sl = s0 + u * 1024;
            int v = 0, d = 0;
            if (sl < n) {
                v = tmp[sl];
                int c = slot_cell[sl];
                int o0 = off[c], o1 = off[c + 1];
                int rank = 0;
#pragma unroll 8
                for (int q = o0; q < o1; q++) rank += tmp[q] < v;
                d = o0 + rank;
            }
            vv[u] = v;
            dst[u] = d;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            pv[u] = pos[vv[u]];
            mv[u] = meta[vv[u]].y;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (s0 + u * 1024 < n) {
                sorted[dst[u]] = vv[u];
                cpos[dst[u]] = make_float4((float)wrap_coord(pv[u].x, b.L[0], b.inv[0], b.periodic),
                                           (float)wrap_coord(pv[u].y, b.L[1], b.inv[1], b.periodic),
                                           (float)wrap_coord(pv[u].z, b.L[2], b.inv[2], b.periodic), __int_as_float(mv[u]));
            }
        }
    }
}

// Verlet list: TPA lanes build the row of atom sorted[t] together.  For each
// (dz, dy) of the neighbour cells, the cells along x form at most three runs
// of consecutive cell indices (split where the periodic wrap is crossed), and
// a run's atoms are consecutive in sorted order, so each pass reads TPA
// consecutive candidate records (FP32 wrapped position + molecule id, one
// coalesced 16-byte load).  The image of a candidate is its cell's periodic
// shift (reach 1 or 2), or the nearest image on axes with one cell.  A
// candidate is kept when it belongs to another molecule and its FP32
// distance is below rlist (1 + 1e-4): FP32 rounding of the wrapped
// coordinates moves distances by far less than that margin, so the list
// holds every pair within rlist; the force kernels apply the exact cutoff.
// Accepted ones are written in candidate order through a ballot prefix
// count, so the row is the same whatever the lane timing.  The kernel also
// clears the per-cell counts for the next rebuild.
template <int TPA>
__global__ __launch_bounds__(256) void k_nlist_build(const float4* __restrict__ cpos, const int* __restrict__ sorted,
                                                     const int* __restrict__ cellid, const int* __restrict__ start,
                                                     int* __restrict__ count_zero, Grid g, Box b, float rl2f,
                                                     int n, int cap, int* __restrict__ nbr, int* __restrict__ nnbr,
                                                     const double4* __restrict__ pos, double4* __restrict__ ref,
                                                     int* __restrict__ flags, const int2* __restrict__ meta) {
    if (flags[0] == 0) return;
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    for (int c = gid; c < g.ncell; c += gridDim.x * blockDim.x) count_zero[c] = 0;
    int t = gid / TPA;
    int sub = gid % TPA;
    int gbase = (threadIdx.x & 63) - sub;
    if (t < n) {
        int i = sorted[t];
        float4 pi = cpos[t];
        int mi = __float_as_int(pi.w);
        int c = cellid[i];
        int cx = c % g.n[0];
        int cy = (c / g.n[0]) % g.n[1];
        int cz = c / (g.n[0] * g.n[1]);
        const float Lx = (float)b.L[0], Ly = (float)b.L[1], Lz = (float)b.L[2];
        const float ix = (float)b.inv[0], iy = (float)b.inv[1], iz = (float)b.inv[2];
        // nearest image on axes searched as a single cell (periodic only)
        const bool nx1 = b.periodic && g.reach[0] == 0, ny1 = b.periodic && g.reach[1] == 0,
                   nz1 = b.periodic && g.reach[2] == 0;
        int count = 0;
        int* row = nbr + (size_t)t * cap;
        // The neighbour runs, in (dz, dy, run) order: run 0 = the cells past
        // the low x edge (shifted by -L), 1 = the cells inside, 2 = past the
        // high edge (+L).  A run's bounds are loaded one run ahead and a
        // pass's candidate record one pass ahead, so the dependent loads of
        // consecutive runs and passes overlap.
        const int ey = 2 * g.reach[1] + 1, nseg = (2 * g.reach[2] + 1) * ey * 3;
        const int x0 = cx - g.reach[0], x1 = cx + g.reach[0];
        struct Run {
            int s0, e0;
            float sx, sy, sz;
        };
        // next run with cells at or after index k (nseg when none)
        auto next_run = [&](int k) {
            while (k < nseg) {
                int run = k % 3;
                if ((run == 0 && x0 < 0) || run == 1 || (run == 2 && x1 >= g.n[0])) return k;
                k++;
            }
            return k;
        };
        auto run_bounds = [&](int k) {
            int run = k % 3, dyi = (k / 3) % ey, dzi = k / (3 * ey);
            int z = cz - g.reach[2] + dzi, y = cy - g.reach[1] + dyi;
            Run r;
            r.sz = 0.f;
            if (z < 0) {
                z += g.n[2];
                r.sz = -Lz;
            } else if (z >= g.n[2]) {
                z -= g.n[2];
                r.sz = Lz;
            }
            r.sy = 0.f;
            if (y < 0) {
                y += g.n[1];
                r.sy = -Ly;
            } else if (y >= g.n[1]) {
                y -= g.n[1];
                r.sy = Ly;
            }
            int xa, xb;
            if (run == 0) {
                xa = x0 + g.n[0];
                xb = g.n[0] - 1;
                r.sx = -Lx;
            } else if (run == 1) {
                xa = x0 < 0 ? 0 : x0;
                xb = x1 >= g.n[0] ? g.n[0] - 1 : x1;
                r.sx = 0.f;
            } else {
                xa = 0;
                xb = x1 - g.n[0];
                r.sx = Lx;
            }
            int rowc = (z * g.n[1] + y) * g.n[0];
            r.s0 = start[rowc + xa];
            r.e0 = start[rowc + xb + 1];
            return r;
        };
        int k = next_run(0);
        Run cur = run_bounds(k);  // k < nseg always: run 1 exists for every (dz, dy)
        // loads are unconditional, with indices clamped into range, so that
        // the compiler can leave them in flight across the work in between
        while (k < nseg) {
            int kn = next_run(k + 1);
            Run nxt = run_bounds(kn < nseg ? kn : k);
            float ddx0 = pi.x - cur.sx, ddy0 = pi.y - cur.sy, ddz0 = pi.z - cur.sz;
            // one pass: TPA consecutive candidates, kept in candidate order by
            // a ballot prefix count
            auto pass = [&](int q, float4 cj, int jj) {
                bool accept = false;
                if (q < cur.e0 && __float_as_int(cj.w) != mi) {
                    float ddx = ddx0 - cj.x, ddy = ddy0 - cj.y, ddz = ddz0 - cj.z;
                    if (nx1) ddx -= rintf(ddx * ix) * Lx;
                    if (ny1) ddy -= rintf(ddy * iy) * Ly;
                    if (nz1) ddz -= rintf(ddz * iz) * Lz;
                    accept = ddx * ddx + ddy * ddy + ddz * ddz < rl2f;
                }
                unsigned long long bal = __ballot(accept);
                unsigned long long gm = TPA == 64 ? bal : ((bal >> gbase) & ((1ull << (TPA & 63)) - 1ull));
                int slot = count + __popcll(gm & ((1ull << sub) - 1ull));
                if (accept && slot < cap) row[slot] = nbr_entry(jj, meta[jj].x);
                count += __popcll(gm);
            };
            // two passes per iteration on alternating registers, each
            // buffer reloaded right after its use, so that a candidate load
            // stays in flight across a whole pass
            auto clampq = [&](int q) { return q < n ? q : n - 1; };
            int q = cur.s0 + sub;
            float4 cA = cpos[clampq(q)];
            int jA = sorted[clampq(q)];
            for (int q0 = cur.s0; q0 < cur.e0; q0 += 2 * TPA) {
                float4 cB = cpos[clampq(q + TPA)];
                int jB = sorted[clampq(q + TPA)];
                pass(q, cA, jA);
                cA = cpos[clampq(q + 2 * TPA)];
                jA = sorted[clampq(q + 2 * TPA)];
                if (q0 + TPA < cur.e0) pass(q + TPA, cB, jB);
                q += 2 * TPA;
            }
            k = kn;
            cur = nxt;
        }
        if (sub == 0) {
            if (count > cap) {
                atomicMax(flags + 1, count);
                count = cap;
            }
            nnbr[t] = count;
            ref[i] = pos[i];
        }
    }
}

// Verlet list, one workgroup per home cell.  The candidate runs of the cell
// (the same (dz, dy, run) segments, in the same order, as k_nlist_build) are
// flattened into one index range; the workgroup stages them in LDS a tile at
// a time, every thread issuing kCellTile / 256 independent loads, with the
// cell's periodic shift already added.  Each wave then scans the tile for the
// home atoms it owns (one atom per wave at a time, 64 candidates a pass).
// Rows come out in the same candidate order as k_nlist_build's, so the
// force sums over them are the same; the acceptance test differs only in
// where the shift is rounded, well inside the 1e-4 margin on rlist.  The
// global loads of a whole tile are in flight together, where the per-atom
// kernel waits on one dependent load per pass; in the scan, a wave tests
// each candidate record (one LDS read) against two home atoms at once,
// without branches.
constexpr int kCellBlock = 512;
constexpr int kCellTile = 2048;  // candidates staged per tile (40 KB of LDS)
constexpr int kCellHome = 64;    // home atoms per chunk
constexpr int kCellMaxSeg = 128; // (2 reach + 1)^2 rows x 3 runs <= 75

// The scan of a staged candidate tile against the home atoms of a workgroup
// (both Verlet builds): each wave tests every staged record against two home
// atoms at once, branch-free, and appends the accepted ones to the home
// atoms' rows (home_row: row index) in candidate order with ballot prefix
// counts; home_cnt carries the row lengths across tiles.
template <bool kMinImage>
__device__ __forceinline__ void tile_scan(const float4* tile, const int* tidx, int tl, const float4* home,
                                          const int* home_row, int* home_cnt, int hn, int* __restrict__ nbr, int cap,
                                          float rl2f, const Box& b, const Grid& g) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    const float Lx = (float)b.L[0], Ly = (float)b.L[1], Lz = (float)b.L[2];
    const float ix = (float)b.inv[0], iy = (float)b.inv[1], iz = (float)b.inv[2];
    const bool nx1 = b.periodic && g.reach[0] == 0, ny1 = b.periodic && g.reach[1] == 0,
               nz1 = b.periodic && g.reach[2] == 0;
    constexpr int NW = kCellBlock / 64;
    auto accept = [&](float4 pi, float4 cj, bool live) {
        float ddx = pi.x - cj.x, ddy = pi.y - cj.y, ddz = pi.z - cj.z;
        if (kMinImage) {
            if (nx1) ddx -= rintf(ddx * ix) * Lx;
            if (ny1) ddy -= rintf(ddy * iy) * Ly;
            if (nz1) ddz -= rintf(ddz * iz) * Lz;
        }
        float d2 = ddx * ddx + ddy * ddy + ddz * ddz;
        return live & (__float_as_int(cj.w) != __float_as_int(pi.w)) & (d2 < rl2f);
    };
    // home atoms h and h + NW of this wave together (the second
    // one repeats the first when absent, and is not written)
    for (int h = wave; h < hn; h += 2 * NW) {
        const bool two = h + NW < hn;
        const int hb = two ? h + NW : h;
        const float4 pa = home[h], pb = home[hb];
        int ca = home_cnt[h], cb = home_cnt[hb];
        int* rowa = nbr + (size_t)home_row[h] * cap;
        int* rowb = nbr + (size_t)home_row[hb] * cap;
        for (int p = 0; p < tl; p += 64) {
            const int e = p + lane;  // < kCellTile: tl <= kCellTile, a multiple of 64
            const float4 cj = tile[e];
            const int jj = tidx[e];
            const bool live = e < tl;
            const bool aa = accept(pa, cj, live), ab = accept(pb, cj, live) & two;
            const unsigned long long ba = __ballot(aa), bb = __ballot(ab);
            const int sa = ca + __popcll(ba & below), sb = cb + __popcll(bb & below);
            if (aa && sa < cap) rowa[sa] = jj;
            if (ab && sb < cap) rowb[sb] = jj;
            ca += __popcll(ba);
            cb += __popcll(bb);
        }
        if (lane == 0) {
            home_cnt[h] = ca;
            if (two) home_cnt[hb] = cb;
        }
    }
}

struct CellsArgs {
    const int2* meta;
    const float4* cpos;
    const int* sorted;
    const int* start;
    int* count_zero;
    Grid g;
    Box b;
    float rl2f;
    int cap;
    int* nbr;
    int* nnbr;
    const double4* pos;
    double4* ref;
    int* flags;
};

// kMinImage: some axis has a single cell (nearest image along it).  Block
// `blk` of `nblk` (cells blk, blk + nblk, ...).
// LDS of one block of the cell build (about 44 KB); a launch that shares its
// blocks with another pass hands both the same buffer.
struct CellsLds {
    float4 tile[kCellTile];
    float4 home[kCellHome];
    int tidx[kCellTile];
    int seg_base[kCellMaxSeg + 1];
    int seg_s0[kCellMaxSeg];
    float seg_sh[kCellMaxSeg][3];
    int home_row[kCellHome];
    int home_cnt[kCellHome];
};

template <bool kMinImage>
__device__ __forceinline__ void nlist_cells_block(const CellsArgs& args, int blk, int nblk, CellsLds& lds) {
    const float4* __restrict__ cpos = args.cpos;
    const int* __restrict__ sorted = args.sorted;
    const int* __restrict__ start = args.start;
    int* __restrict__ count_zero = args.count_zero;
    const Grid& g = args.g;
    const Box& b = args.b;
    const float rl2f = args.rl2f;
    const int cap = args.cap;
    int* __restrict__ nbr = args.nbr;
    int* __restrict__ nnbr = args.nnbr;
    const double4* __restrict__ pos = args.pos;
    double4* __restrict__ ref = args.ref;
    int* __restrict__ flags = args.flags;
    if (flags[0] == 0) return;
    float4* tile = lds.tile;
    int* tidx = lds.tidx;
    int* seg_base = lds.seg_base;
    int* seg_s0 = lds.seg_s0;
    float(*seg_sh)[3] = lds.seg_sh;
    float4* home = lds.home;
    int* home_row = lds.home_row;
    int* home_cnt = lds.home_cnt;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int c = blk * kCellBlock + tid; c < g.ncell; c += nblk * kCellBlock) count_zero[c] = 0;
    const float Lx = (float)b.L[0], Ly = (float)b.L[1], Lz = (float)b.L[2];
    const int ey = 2 * g.reach[1] + 1, nseg = (2 * g.reach[2] + 1) * ey * 3;
    for (int c = blk; c < g.ncell; c += nblk) {
        const int hs = start[c], H = start[c + 1] - hs;
        if (H == 0) continue;  // uniform over the workgroup
        const int cx = c % g.n[0], cy = (c / g.n[0]) % g.n[1], cz = c / (g.n[0] * g.n[1]);
        const int x0 = cx - g.reach[0], x1 = cx + g.reach[0];
        // segment table: wave 0, lane l owns segments 2l and 2l + 1
        if (wave == 0) {
            int len[2];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                int k = 2 * lane + u;
                len[u] = 0;
                int run = k % 3;
                bool exists = k < nseg && (run == 1 || (run == 0 && x0 < 0) || (run == 2 && x1 >= g.n[0]));
                if (exists) {
                    int dyi = (k / 3) % ey, dzi = k / (3 * ey);
                    int z = cz - g.reach[2] + dzi, y = cy - g.reach[1] + dyi;
                    float sz = 0.f, sy = 0.f, sx = 0.f;
                    if (z < 0) {
                        z += g.n[2];
                        sz = -Lz;
                    } else if (z >= g.n[2]) {
                        z -= g.n[2];
                        sz = Lz;
                    }
                    if (y < 0) {
                        y += g.n[1];
                        sy = -Ly;
                    } else if (y >= g.n[1]) {
                        y -= g.n[1];
                        sy = Ly;
                    }
                    int xa, xb;
                    if (run == 0) {
                        xa = x0 + g.n[0];
                        xb = g.n[0] - 1;
                        sx = -Lx;
                    } else if (run == 1) {
                        xa = x0 < 0 ? 0 : x0;
                        xb = x1 >= g.n[0] ? g.n[0] - 1 : x1;
                    } else {
                        xa = 0;
                        xb = x1 - g.n[0];
                        sx = Lx;
                    }
                    int rowc = (z * g.n[1] + y) * g.n[0];
                    int s0 = start[rowc + xa];
                    len[u] = start[rowc + xb + 1] - s0;
                    seg_s0[k] = s0;
                    seg_sh[k][0] = sx;
                    seg_sh[k][1] = sy;
                    seg_sh[k][2] = sz;
                }
            }
            int pair = len[0] + len[1], incl = pair;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                int y = __shfl_up(incl, off, 64);
                if (lane >= off) incl += y;
            }
            seg_base[2 * lane] = incl - pair;
            seg_base[2 * lane + 1] = incl - pair + len[0];
            if (lane == 63) seg_base[kCellMaxSeg] = incl;
        }
        __syncthreads();
        const int M = seg_base[kCellMaxSeg];
        for (int h0 = 0; h0 < H; h0 += kCellHome) {
            const int hn = min(kCellHome, H - h0);
            if (tid < hn) {
                home[tid] = cpos[hs + h0 + tid];
                home_row[tid] = hs + h0 + tid;
                home_cnt[tid] = 0;
            }
            for (int m0 = 0; m0 < M; m0 += kCellTile) {
                const int tl = min(kCellTile, M - m0);
                constexpr int R = kCellTile / kCellBlock;
                int qv[R], kv[R];
                // largest k with seg_base[k] <= mm: entries past the last
                // segment hold M > mm, so no bound test is needed
#pragma unroll
                for (int r = 0; r < R; r++) {
                    int m = m0 + tid + r * kCellBlock;
                    int mm = m < M ? m : M - 1;
                    int k = 0;
#pragma unroll
                    for (int step = kCellMaxSeg / 2; step > 0; step >>= 1) k = seg_base[k + step] <= mm ? k + step : k;
                    kv[r] = k;
                    qv[r] = seg_s0[k] + (mm - seg_base[k]);
                }
                float4 cv[R];
                int jv[R], kv2[R];
#pragma unroll
                for (int r = 0; r < R; r++) {
                    cv[r] = cpos[qv[r]];
                    jv[r] = sorted[qv[r]];
                }
#pragma unroll
                for (int r = 0; r < R; r++) kv2[r] = args.meta[jv[r]].x;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    int e = tid + r * kCellBlock;
                    if (e < tl) {
                        int k = kv[r];
                        tile[e] = make_float4(cv[r].x + seg_sh[k][0], cv[r].y + seg_sh[k][1],
                                              cv[r].z + seg_sh[k][2], cv[r].w);
                        tidx[e] = nbr_entry(jv[r], kv2[r]);
                    }
                }
                __syncthreads();
                tile_scan<kMinImage>(tile, tidx, tl, home, home_row, home_cnt, hn, nbr, cap, rl2f, b, g);
                __syncthreads();
            }
            if (tid < hn) {
                int t = hs + h0 + tid;
                int cnt = home_cnt[tid];
                if (cnt > cap) {
                    atomicMax(flags + 1, cnt);
                    cnt = cap;
                }
                nnbr[t] = cnt;
                int i = sorted[t];
                ref[i] = pos[i];
            }
            __syncthreads();
        }
    }
}

template <bool kMinImage>
__global__ __launch_bounds__(kCellBlock) void k_nlist_cells(CellsArgs ca) {
    __shared__ CellsLds lds;
    nlist_cells_block<kMinImage>(ca, blockIdx.x, gridDim.x, lds);
}

// ---- direct Verlet build (systems up to 32k atoms) --------------------------------
// The same rows as the cell build without a sorted cell list: workgroup c
// reads every atom's cell tag (CellTag, written when the position was) and
// takes, in atom-index order, the atoms of its own cell as home atoms and
// those of the cells within reach as candidates, with the periodic shift of
// the image they stand for (a cell index past the reach on one side is a
// neighbour on the other).  Candidates go to an LDS list (windows of
// kDirectCand, re-read past that), then in tiles through tile_scan; rows are
// indexed by atom (sorted[i] = i), so no launch has to sort anything and a
// step without a rebuild pays no cell-list launch.  A cell's home atoms are
// dealt to `split` workgroups (home rank r to slice r % split), each scanning
// the whole candidate list for its slice, so that a grid of few cells still
// fills the device; every row is still built by one workgroup, in candidate
// order, so the rows do not depend on the split.
constexpr int kDirectCand = 4096;
constexpr int kDirectHome = 256;
constexpr int kDirectBatch = 8;  // scan rounds (of one atom per thread) per pair of barriers

struct DirectArgs {
    const double4* pos;
    const int2* meta;
    const int* tag;
    int n;
    Grid g;
    Box b;
    float rl2f;
    int cap;
    int* nbr;
    int* nnbr;
    double4* ref;
    int* sorted;
    int* flags;
    int split;  // workgroups per cell (1..8)
};

struct DirectLds {
    float4 tile[kCellTile];
    float4 home[kDirectHome];
    int tidx[kCellTile];
    int cand[kDirectCand];  // atom | shift code << 24 (2 bits per axis: 0, +L, -L)
    int home_row[kDirectHome];
    int home_cnt[kDirectHome];
    int wc[kDirectBatch][kCellBlock / 64], wh[kDirectBatch][kCellBlock / 64];
};

__device__ __forceinline__ float direct_shift(int code, float L) { return code == 1 ? L : (code == 2 ? -L : 0.f); }

template <bool kMinImage>
__device__ __forceinline__ void nlist_direct_block(const DirectArgs& a, int blk, int nblk, DirectLds& lds) {
    if (a.flags[0] == 0) return;
    constexpr int NW = kCellBlock / 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    const Grid& g = a.g;
    const Box& b = a.b;
    const int n = a.n;
    const float Lx = (float)b.L[0], Ly = (float)b.L[1], Lz = (float)b.L[2];
    const int split = a.split;
    for (int cs = blk; cs < g.ncell * split; cs += nblk) {
        const int c = cs / split, sl = cs - c * split;
        const int hx = c % g.n[0], hy = (c / g.n[0]) % g.n[1], hz = c / (g.n[0] * g.n[1]);
        int hwin = 0, Hs = 0;
        do {
            int cwin = 0, M = 0;
            do {
                // scan: block-wide ranks of candidates and home atoms in index
                // order, kDirectBatch rounds per pair of barriers
                int basec = 0, baseh = 0;
                for (int i0 = 0; i0 < n; i0 += kDirectBatch * kCellBlock) {
                    int tv[kDirectBatch];
#pragma unroll
                    for (int u = 0; u < kDirectBatch; u++) {
                        const int i = i0 + u * kCellBlock + tid;
                        tv[u] = a.tag[i < n ? i : n - 1];
                    }
                    unsigned long long bc[kDirectBatch], bh[kDirectBatch];
                    int code[kDirectBatch];
#pragma unroll
                    for (int u = 0; u < kDirectBatch; u++) {
                        const int i = i0 + u * kCellBlock + tid;
                        const int t = tv[u];
                        int dx = (t & 1023) - hx, dy = ((t >> 10) & 1023) - hy, dz = (t >> 20) - hz;
                        int cd = 0;
                        if (dx > g.reach[0]) {
                            dx -= g.n[0];
                            cd |= 2;
                        } else if (dx < -g.reach[0]) {
                            dx += g.n[0];
                            cd |= 1;
                        }
                        if (dy > g.reach[1]) {
                            dy -= g.n[1];
                            cd |= 2 << 2;
                        } else if (dy < -g.reach[1]) {
                            dy += g.n[1];
                            cd |= 1 << 2;
                        }
                        if (dz > g.reach[2]) {
                            dz -= g.n[2];
                            cd |= 2 << 4;
                        } else if (dz < -g.reach[2]) {
                            dz += g.n[2];
                            cd |= 1 << 4;
                        }
                        const bool live = i < n;
                        const bool cand =
                            live && abs(dx) <= g.reach[0] && abs(dy) <= g.reach[1] && abs(dz) <= g.reach[2];
                        const bool home = live && (dx | dy | dz | cd) == 0;
                        code[u] = cd;
                        bc[u] = __ballot(cand);
                        bh[u] = __ballot(home);
                    }
                    if (lane == 0)
#pragma unroll
                        for (int u = 0; u < kDirectBatch; u++) {
                            lds.wc[u][wave] = __popcll(bc[u]);
                            lds.wh[u][wave] = __popcll(bh[u]);
                        }
                    __syncthreads();
#pragma unroll
                    for (int u = 0; u < kDirectBatch; u++) {
                        const int i = i0 + u * kCellBlock + tid;
                        int offc = 0, offh = 0, totc = 0, toth = 0;
#pragma unroll
                        for (int w = 0; w < NW; w++) {
                            const int xc = lds.wc[u][w], xh = lds.wh[u][w];
                            if (w < wave) {
                                offc += xc;
                                offh += xh;
                            }
                            totc += xc;
                            toth += xh;
                        }
                        const bool cand = (bc[u] >> lane) & 1ull, home = (bh[u] >> lane) & 1ull;
                        const int rc = basec + offc + __popcll(bc[u] & below) - cwin;
                        if (cand && rc >= 0 && rc < kDirectCand) lds.cand[rc] = i | (code[u] << 24);
                        const int rg = baseh + offh + __popcll(bh[u] & below);  // rank in the cell
                        const int rh = rg / split - hwin;                       // rank in the slice window
                        if (cwin == 0 && home && rg - (rg / split) * split == sl && rh >= 0 && rh < kDirectHome)
                            lds.home_row[rh] = i;
                        basec += totc;
                        baseh += toth;
                    }
                    __syncthreads();  // wc / wh are rewritten by the next batch
                }
                M = basec;
                Hs = (baseh - sl + split - 1) / split;  // home ranks r < baseh with r % split == sl
                const int hn = min(kDirectHome, Hs - hwin);
                if (hn <= 0) break;  // uniform: an empty cell or slice
                if (cwin == 0)
                    for (int h = tid; h < hn; h += kCellBlock) {
                        const int i = lds.home_row[h];
                        const double4 p = a.pos[i];
                        lds.home[h] = make_float4((float)wrap_coord(p.x, b.L[0], b.inv[0], b.periodic),
                                                  (float)wrap_coord(p.y, b.L[1], b.inv[1], b.periodic),
                                                  (float)wrap_coord(p.z, b.L[2], b.inv[2], b.periodic),
                                                  __int_as_float(a.meta[i].y));
                        lds.home_cnt[h] = 0;
                    }
                const int mw = min(kDirectCand, M - cwin);
                for (int t0 = 0; t0 < mw; t0 += kCellTile) {
                    const int tl = min(kCellTile, mw - t0);
                    __syncthreads();  // the list and homes are written; the previous tile is scanned
                    constexpr int R = kCellTile / kCellBlock;
                    double4 pv[R];
                    int2 mv[R];
                    int vv[R];
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const int e = tid + r * kCellBlock;
                        vv[r] = e < tl ? lds.cand[t0 + e] : lds.cand[t0];
                        const int i = vv[r] & 0xffffff;
                        pv[r] = a.pos[i];
                        mv[r] = a.meta[i];
                    }
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const int e = tid + r * kCellBlock;
                        if (e < tl) {
                            const int code = vv[r] >> 24;
                            lds.tile[e] = make_float4(
                                (float)wrap_coord(pv[r].x, b.L[0], b.inv[0], b.periodic) + direct_shift(code & 3, Lx),
                                (float)wrap_coord(pv[r].y, b.L[1], b.inv[1], b.periodic) + direct_shift((code >> 2) & 3, Ly),
                                (float)wrap_coord(pv[r].z, b.L[2], b.inv[2], b.periodic) + direct_shift(code >> 4, Lz),
                                __int_as_float(mv[r].y));
                            lds.tidx[e] = nbr_entry(vv[r] & 0xffffff, mv[r].x);
                        }
                    }
                    __syncthreads();
                    tile_scan<kMinImage>(lds.tile, lds.tidx, tl, lds.home, lds.home_row, lds.home_cnt, hn, a.nbr, a.cap,
                                         a.rl2f, b, g);
                }
                __syncthreads();
                cwin += kDirectCand;
            } while (cwin < M);
            const int hn = min(kDirectHome, Hs - hwin);
            for (int h = tid; h < hn; h += kCellBlock) {
                const int i = lds.home_row[h];
                int cnt = lds.home_cnt[h];
                if (cnt > a.cap) {
                    atomicMax(a.flags + 1, cnt);
                    cnt = a.cap;
                }
                a.nnbr[i] = cnt;
                a.sorted[i] = i;
                a.ref[i] = a.pos[i];
            }
            __syncthreads();
            hwin += kDirectHome;
        } while (hwin < Hs);
    }
}

template <bool kMinImage>
__global__ __launch_bounds__(kCellBlock) void k_nlist_direct(DirectArgs da) {
    __shared__ DirectLds lds;
    nlist_direct_block<kMinImage>(da, blockIdx.x, gridDim.x, lds);
}



// ---------------------------------------------------------------------------------
// Intra-molecular terms: one thread per term (same-molecule pair with its bond
// path, bond, angle or dihedral), compute.rs:60-97, energy.rs:88-138,
// ewald.rs:387-428 (excluded-pair correction), cells.rs:335-411.  Each term
// writes its force on each of its atoms to a slot reserved for that
// (term, atom); the consumer adds an atom's slots in slot order, so the sums
// are deterministic without atomics.  Terms are sorted by type on the host so
// that waves do not diverge.
// ---------------------------------------------------------------------------------
enum : int { TERM_PAIR = 0, TERM_BOND = 1, TERM_ANGLE = 2, TERM_DIHEDRAL = 3 };

struct Term {
    int4 atoms;  // pair/bond: x, y; angle: x, y, z; dihedral: x, y, z, w
    int4 slots;  // contribution slot of each atom
    int type;    // TERM_*
    int param;   // pair: bond path; bond/angle/dihedral: parameter index
};

struct IntraArgs {
    const double4* pos;
    const int2* meta;
    int nterms;
    const Term* terms;
    const PairParam* table;
    int nkinds;
    const int* pair_restr;  // restriction per kind pair
    const double* pair_s14; // scale14 per kind pair
    const BondedParam* bparam;
    const BondedParam* aparam;
    const BondedParam* dparam;
    const double* tabs;  // TableComputation tables of the pair potentials
    CoulParam coul;
    int coul_restr;
    double coul_s14;
    Box box;
    double4* contrib;  // slot contributions
    double* epart;     // per block: pairs, coul, bonds, angles, dihedrals, w6
};

// One block (256 threads) of intra-molecular terms, block index `blk`: run as
// its own kernel or as the extra blocks of a pair-kernel launch.
template <bool kEnergy>
__device__ __forceinline__ void intra_block(const IntraArgs& a, int blk) {
    __shared__ double red[4][11];
    int t = blk * blockDim.x + threadIdx.x;
    double acc[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (t < a.nterms) {
        Term tm = a.terms[t];
        if (tm.type == TERM_PAIR) {
            int i = tm.atoms.x, j = tm.atoms.y, path = tm.param;
            double4 pi = a.pos[i], pj = a.pos[j];
            double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            min_image(a.box, dx, dy, dz);
            double r = sqrt(norm2_ordered(dx, dy, dz));
            double inv_r = 1.0 / r;
            int tk = a.meta[i].x * a.nkinds + a.meta[j].x;
            const PairParam& pp = a.table[tk];
            double fs = 0.0;
            if (pp.pot != POT_NONE) {
                bool excl;
                double scaling;
                restriction_info(a.pair_restr[tk], a.pair_s14[tk], path, excl, scaling);
                if (!excl && r < pp.rc && pp.pot != POT_NULL) {
                    double e, f;
                    pot_eval_any(pp, a.tabs, r, inv_r, e, f);
                    fs += scaling * f;
                    acc[0] += scaling * (e - pp.shift);
                }
            }
            double qiqj = pi.w * pj.w;
            if (a.coul.kind != COUL_NONE && qiqj != 0.0) {
                bool excl;
                double scaling;
                restriction_info(a.coul_restr, a.coul_s14, path, excl, scaling);
                if (r <= a.coul.rc) {
                    double e, f;
                    if (a.coul.kind == COUL_EWALD) {
                        ewald_real(a.coul.alpha, qiqj, r, inv_r, excl, e, f);
                        fs += f;
                        acc[1] += e;
                    } else if (!excl) {
                        wolf_pair(a.coul, qiqj, r, inv_r, e, f);
                        fs += scaling * f;
                        acc[1] += scaling * e;
                    }
                }
            }
            a.contrib[tm.slots.x] = make_double4(fs * dx, fs * dy, fs * dz, 0.0);
            a.contrib[tm.slots.y] = make_double4(-fs * dx, -fs * dy, -fs * dz, 0.0);
            acc[5] += fs * dx * dx;
            acc[6] += fs * dy * dy;
            acc[7] += fs * dz * dz;
            acc[8] += fs * dx * dy;
            acc[9] += fs * dx * dz;
            acc[10] += fs * dy * dz;
        } else if (tm.type == TERM_BOND) {
            // compute.rs:63-73; virial compute.rs:231-239
            const BondedParam& bp = a.bparam[tm.param];
            double fx = 0, fy = 0, fz = 0;
            if (bp.pot != POT_NONE) {
                double4 pi = a.pos[tm.atoms.x], pj = a.pos[tm.atoms.y];
                double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                min_image(a.box, dx, dy, dz);
                double r = sqrt(norm2_ordered(dx, dy, dz));
                double e, f;
                pot_eval_any(bp, nullptr, r, 1.0 / r, e, f);
                fx = f * dx;
                fy = f * dy;
                fz = f * dz;
                acc[2] += e;
                acc[5] += f * dx * dx;
                acc[6] += f * dy * dy;
                acc[7] += f * dz * dz;
                acc[8] += f * dx * dy;
                acc[9] += f * dx * dz;
                acc[10] += f * dy * dz;
            }
            a.contrib[tm.slots.x] = make_double4(fx, fy, fz, 0.0);
            a.contrib[tm.slots.y] = make_double4(-fx, -fy, -fz, 0.0);
        } else if (tm.type == TERM_ANGLE) {
            // cells.rs:335-359
            const BondedParam& ap = a.aparam[tm.param];
            double4 c0 = make_double4(0, 0, 0, 0), c1 = c0, c2 = c0;
            if (ap.pot != POT_NONE) {
                double4 p1 = a.pos[tm.atoms.x], p2 = a.pos[tm.atoms.y], p3 = a.pos[tm.atoms.z];
                double ax = p1.x - p2.x, ay = p1.y - p2.y, az = p1.z - p2.z;
                double bx = p3.x - p2.x, by = p3.y - p2.y, bz = p3.z - p2.z;
                min_image(a.box, ax, ay, az);
                min_image(a.box, bx, by, bz);
                double na = sqrt(norm2_ordered(ax, ay, az)), nb = sqrt(norm2_ordered(bx, by, bz));
                double ux = ax / na, uy = ay / na, uz = az / na;
                double vx = bx / nb, vy = by / nb, vz = bz / nb;
                double cs = ux * vx + uy * vy + uz * vz;
                double sin_inv = 1.0 / sqrt(1.0 - cs * cs);
                double d1x = sin_inv * (cs * ux - vx) / na, d1y = sin_inv * (cs * uy - vy) / na,
                       d1z = sin_inv * (cs * uz - vz) / na;
                double d3x = sin_inv * (cs * vx - ux) / nb, d3y = sin_inv * (cs * vy - uy) / nb,
                       d3z = sin_inv * (cs * vz - uz) / nb;
                double theta = acos(cs);
                double e, f;
                // with inv_r = 1 pot_eval returns the plain -dV/dtheta
                pot_eval_any(ap, nullptr, theta, 1.0, e, f);
                c0 = make_double4(f * d1x, f * d1y, f * d1z, 0.0);
                c1 = make_double4(-f * (d1x + d3x), -f * (d1y + d3y), -f * (d1z + d3z), 0.0);
                c2 = make_double4(f * d3x, f * d3y, f * d3z, 0.0);
                acc[3] += e;
            }
            a.contrib[tm.slots.x] = c0;
            a.contrib[tm.slots.y] = c1;
            a.contrib[tm.slots.z] = c2;
        } else {
            // cells.rs:379-411
            const BondedParam& dp = a.dparam[tm.param];
            double4 c0 = make_double4(0, 0, 0, 0), c1 = c0, c2 = c0, c3 = c0;
            if (dp.pot != POT_NONE) {
                double4 q1 = a.pos[tm.atoms.x], q2 = a.pos[tm.atoms.y], q3 = a.pos[tm.atoms.z],
                        q4 = a.pos[tm.atoms.w];
                double r12x = q2.x - q1.x, r12y = q2.y - q1.y, r12z = q2.z - q1.z;
                double r23x = q3.x - q2.x, r23y = q3.y - q2.y, r23z = q3.z - q2.z;
                double r34x = q4.x - q3.x, r34y = q4.y - q3.y, r34z = q4.z - q3.z;
                min_image(a.box, r12x, r12y, r12z);
                min_image(a.box, r23x, r23y, r23z);
                min_image(a.box, r34x, r34y, r34z);
                double ux = r12y * r23z - r12z * r23y, uy = r12z * r23x - r12x * r23z,
                       uz = r12x * r23y - r12y * r23x;
                double vx = r23y * r34z - r23z * r34y, vy = r23z * r34x - r23x * r34z,
                       vz = r23x * r34y - r23y * r34x;
                double u2 = ux * ux + uy * uy + uz * uz, v2 = vx * vx + vy * vy + vz * vz;
                double r23_2 = r23x * r23x + r23y * r23y + r23z * r23z;
                double r23n = sqrt(r23_2);
                double d1x = (-r23n / u2) * ux, d1y = (-r23n / u2) * uy, d1z = (-r23n / u2) * uz;
                double d4x = (r23n / v2) * vx, d4y = (r23n / v2) * vy, d4z = (r23n / v2) * vz;
                double r23_r34 = r23x * r34x + r23y * r34y + r23z * r34z;
                double r12_r23 = r12x * r23x + r12y * r23y + r12z * r23z;
                double k1 = -r12_r23 / r23_2 - 1.0, k2 = r23_r34 / r23_2;
                double k3 = -r23_r34 / r23_2 - 1.0, k4 = r12_r23 / r23_2;
                double d2x = k1 * d1x + k2 * d4x, d2y = k1 * d1y + k2 * d4y, d2z = k1 * d1z + k2 * d4z;
                double d3x = k3 * d4x + k4 * d1x, d3y = k3 * d4y + k4 * d1y, d3z = k3 * d4z + k4 * d1z;
                double phi = atan2(r23n * (vx * r12x + vy * r12y + vz * r12z), ux * vx + uy * vy + uz * vz);
                double e, f;
                pot_eval_any(dp, nullptr, phi, 1.0, e, f);
                c0 = make_double4(f * d1x, f * d1y, f * d1z, 0.0);
                c1 = make_double4(f * d2x, f * d2y, f * d2z, 0.0);
                c2 = make_double4(f * d3x, f * d3y, f * d3z, 0.0);
                c3 = make_double4(f * d4x, f * d4y, f * d4z, 0.0);
                acc[4] += e;
            }
            a.contrib[tm.slots.x] = c0;
            a.contrib[tm.slots.y] = c1;
            a.contrib[tm.slots.z] = c2;
            a.contrib[tm.slots.w] = c3;
        }
    }
    if (kEnergy) {
        int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        for (int q = 0; q < 11; q++) {
            double sum = wave_sum(acc[q]);
            if (lane == 0) red[wid][q] = sum;
        }
        __syncthreads();
        if (threadIdx.x < 11) {
            double sum = 0;
            for (int w = 0; w < 4; w++) sum += red[w][threadIdx.x];
            a.epart[blk * 11 + threadIdx.x] = sum;
        }
    }
}

template <bool kEnergy>
__global__ __launch_bounds__(256) void k_intra_terms(IntraArgs a) {
    intra_block<kEnergy>(a, blockIdx.x);
}

// Per k-vector record of the force kernels: rho(k) (written by
// k_rho_reduce) and g = 2 factor_k, gz = g * ikz.  The reference's force
// factor field_k = 2 factor_k k (ewald.rs:712-714) has k = (gx ikx, gy iky,
// gz' ikz) with gx = 2 pi / Lx..., so along a row only the z component
// varies: the kernels sum A = sum_k s_k g and B = sum_k s_k gz and apply
// (gx ikx A, gy iky A, gz' B) once per row.
struct alignas(32) KRec {
    double re, im, g, gz;
};
struct alignas(16) KRec32 {
    float re, im, g, gz;
};

// rho(k) = sum over chunks, in a fixed order: thread (g, kl) of the block
// sums the chunks c = g mod 16 for k = 16 * block + kl, then 16 threads add
// the 16 group sums in order.  With kEnergy the block writes its energy and
// virial partials: e = factor |rho|^2 and w = |rho|^2 * virial_k
// (6 components).  grid: ceil(nk / kRhoReduceK), block 256.
constexpr int kRhoReduceK = 16;

struct RhoReduceArgs {
    const double2* partial;
    int nchunks, nslot;
    const int* kslot;
    int nk;
    double2* rho;
    KRec* rec;
    KRec32* rec32;
    const double* efac;
    const double* vfac;
    double* epart;
};

// One block (256 threads) of the reduction, block index `blk`: run as its
// own kernel or as the last blocks of a pair-kernel launch.
template <bool kEnergy>
__device__ __forceinline__ void rho_reduce_block(const RhoReduceArgs& a, int blk) {
    const double2* __restrict__ partial = a.partial;
    const int nchunks = a.nchunks, nslot = a.nslot, nk = a.nk;
    constexpr int G = 256 / kRhoReduceK;
    __shared__ double2 part[G][kRhoReduceK];
    int g = threadIdx.x / kRhoReduceK, kl = threadIdx.x % kRhoReduceK;
    int k = blk * kRhoReduceK + kl;
    double re = 0.0, im = 0.0;
    if (k < nk) {
        const int slot = a.kslot[k];
#pragma unroll 4
        for (int c = g; c < nchunks; c += G) {  // loads in flight together, sums in order
            double2 p = partial[(size_t)c * nslot + slot];
            re += p.x;
            im += p.y;
        }
    }
    part[g][kl] = make_double2(re, im);
    __syncthreads();
    if (threadIdx.x >= 64) return;
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    if (threadIdx.x < kRhoReduceK && k < nk) {
        re = 0.0;
        im = 0.0;
        for (int q = 0; q < G; q++) {
            re += part[q][kl].x;
            im += part[q][kl].y;
        }
        a.rho[k] = make_double2(re, im);
        if (a.rec) {
            a.rec[k].re = re;
            a.rec[k].im = im;
        }
        if (a.rec32) {
            a.rec32[k].re = (float)re;
            a.rec32[k].im = (float)im;
        }
        if (kEnergy) {
            double n2 = re * re + im * im;
            v[0] = a.efac[k] * n2;
            for (int x = 0; x < 6; x++) v[1 + x] = n2 * a.vfac[(size_t)x * nk + k];
        }
    }
    if (kEnergy) {
        for (int x = 0; x < 7; x++) {
            double sum = wave_sum(v[x]);
            if (threadIdx.x == 0) a.epart[blk * 7 + x] = sum;
        }
    }
}

template <bool kEnergy>
__global__ __launch_bounds__(256) void k_rho_reduce(RhoReduceArgs a) {
    rho_reduce_block<kEnergy>(a, blockIdx.x);
}

// ---------------------------------------------------------------------------------
// Non-bonded pair kernel: inter-molecular pairs, Lennard-Jones family + Ewald
// real space or Wolf, one lane per atom i over its neighbour list.
// Forces are written for every i (full list, both directions), so no atomics
// are needed and the per-atom sums are deterministic.  Energies and the
// virial are accumulated for j > i only, the reference's i < j pair set.
// ---------------------------------------------------------------------------------
struct PairArgs {
    const double4* pos;
    const int2* meta;
    const int* sorted;
    const int* nbr;
    const int* nnbr;
    int cap;
    const PairParam* table;
    int nkinds;
    CoulParam coul;
    Box box;
    int n;
    double4* force;
    double* epart;  // per block: e_pairs, e_coul, w_xx, w_yy, w_zz, w_xy, w_xz, w_yz
    int* rebuild;   // neighbour-list rebuild flag, cleared here (see k_nlist_check)
    // blocks [0, pair_blocks) evaluate pairs; the next intra_blocks run the
    // intra-molecular terms and the rest the rho(k) reduction
    int pair_blocks, intra_blocks;
    IntraArgs intra;
    RhoReduceArgs rr;
    // erfcx tables of the Ewald / Wolf real-space erfc (erfcx_tab), or null
    // for the library erfc
    const double* ecx64;
    const float* ecx32;
    double ecx_inv_h;
    const double* tabs;  // TableComputation tables (general instance)
    const float4* frac;  // positions as frac_pos (mixed instance)
    // the real-space force factor h(r) = erfc(alpha r) + g0 r exp(-alpha^2 r^2)
    // in FP32 pieces over [0, rc] (mixed instance, Ewald and Wolf), or null
    const float* hf32;
    double hf_inv_h;
};

constexpr int kMaxKinds = 16;
constexpr int kPairBlock = 256;

// Blocks are dealt round-robin to the eight XCDs (block b runs on XCD b %
// 8).  The pair kernels take block b as logical block (b % 8) * (P / 8) +
// b / 8 of P (a multiple of 8), so that each XCD works on one contiguous
// eighth of the atoms: in a spatially ordered system its L2 then holds the
// positions of one region and its neighbours (a 1M-atom Lennard-Jones box:
// about a fifth of the 16 MB of positions per 4 MB L2) instead of the
// whole box.
__device__ __forceinline__ int xcd_contiguous(int b, int P) { return (b & 7) * (P >> 3) + (b >> 3); }

// Coulomb variant of a pair kernel instance, fixed per system: none, Ewald
// real space (erfc), Ewald for inter-molecular pairs that the restriction
// excludes (the -erf removal term only), Wolf.  A Wolf system whose
// restriction excludes the pairs uses CK_NONE.  One variant per instance
// keeps the registers of the unused paths out of the kernel.
enum : int { CK_NONE = 0, CK_EWALD = 1, CK_EWALD_ERF = 2, CK_WOLF = 3 };

__host__ __forceinline__ int coulomb_variant(const CoulParam& c) {
    if (c.kind == COUL_EWALD) return c.inter_excl ? CK_EWALD_ERF : CK_EWALD;
    if (c.kind == COUL_WOLF) return c.inter_excl ? CK_NONE : CK_WOLF;
    return CK_NONE;
}

// TPA lanes share one atom i and stride through its neighbour row; the
// group's partial forces are combined with xor-shuffles (TPA divides 64).
// Held to 128 VGPRs (four waves per SIMD instead of three): the kernel waits
// on its gathers most of the time, and the extra resident wave outweighs the
// few spilled registers (SPC/E FP64 pairs 44 -> 40 us per launch).
// kGen: the general instance for systems with a pair potential other than
// LJ/Harmonic/Null or a TableComputation (any kind, FP64, both forms); the
// fast instances stage only the LJ/Harmonic fields.
template <bool kEnergy, int TPA, int CK, bool kGen = false>
__global__ __launch_bounds__(kPairBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_pairs(PairArgs a) {
    using TabEntry = typename std::conditional<kGen, PairParam, PairFast>::type;
    __shared__ TabEntry tab[kMaxKinds * kMaxKinds];
    __shared__ double red[kPairBlock / 64][8];
    constexpr bool kTab = CK == CK_EWALD || CK == CK_WOLF;
    __shared__ double ecx[kTab ? kErfcxPieces * (kErfcxDeg64 + 1) : 1];
    if ((int)blockIdx.x >= a.pair_blocks) {  // uniform over the block
        const int b = blockIdx.x - a.pair_blocks;
        if (b < a.intra_blocks)
            intra_block<kEnergy>(a.intra, b);
        else
            rho_reduce_block<kEnergy>(a.rr, b - a.intra_blocks);
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.rebuild) *a.rebuild = 0;
    int nk2 = a.nkinds * a.nkinds;
    for (int x = threadIdx.x; x < nk2; x += blockDim.x) {
        if constexpr (kGen)
            tab[x] = a.table[x];
        else
            tab[x] = pair_fast(a.table[x]);
    }
    const bool use_tab = kTab && a.ecx64 != nullptr;  // uniform
    if (use_tab)
        for (int x = threadIdx.x; x < kErfcxPieces * (kErfcxDeg64 + 1); x += blockDim.x) ecx[x] = a.ecx64[x];
    __syncthreads();

    // XCD-contiguous order of the pair blocks (the launch has a multiple of 8)
    int g = xcd_contiguous(blockIdx.x, a.pair_blocks) * blockDim.x + threadIdx.x;
    int t = g / TPA;
    int sub = g % TPA;
    double fx = 0, fy = 0, fz = 0;
    double e_pair = 0, e_coul = 0;
    double wxx = 0, wyy = 0, wzz = 0, wxy = 0, wxz = 0, wyz = 0;
    int i = -1;
    if (t < a.n) {
        i = a.sorted[t];
        double4 pi = a.pos[i];
        int ki = a.meta[i].x;
        double qi = pi.w;
        int nn = a.nnbr[t];
        const CoulParam& c = a.coul;
        const int* row = a.nbr + (size_t)t * a.cap;
        // Software pipeline over the row: the row entries two passes ahead
        // and the next pass's neighbour positions are loaded while the
        // current pair is evaluated, unrolled by two on alternating
        // registers (A, B); an entry carries the neighbour's kind.  Entries
        // past the row end read a clamped index and stand for atom i
        // itself, never evaluated.
        auto jat = [&](int k) {
            int v = row[min(k, max(nn - 1, 0))];
            return k < nn ? v : nbr_entry(i, ki);
        };
        auto pair = [&](int j, const double4& pj, int kj) {
            double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            min_image(a.box, dx, dy, dz);
            double r2 = norm2_ordered(dx, dy, dz);
            double r = sqrt(r2);
            double inv_r = 1.0 / r;
            double fs = 0.0;
            double e_p = 0.0, e_c = 0.0;
            const TabEntry& pp = tab[ki * a.nkinds + kj];
            if (pp.pot > POT_NULL && !pp.inter_excl && r < pp.rc) {
                double e, f;
                if constexpr (kGen)
                    pot_eval_any(pp, a.tabs, r, inv_r, e, f);
                else
                    pot_eval(pp.pot, pp.p0, pp.p1, r, inv_r, e, f);
                fs += f;
                e_p = e - pp.shift;
            }
            double qiqj = qi * pj.w;
            if (CK != CK_NONE && qiqj != 0.0 && r <= c.rc) {
                double e, f;
                if (kTab && use_tab) {
                    // ewald.rs:387-428 / wolf.rs:90-117 with erfc = exp(-x^2) erfcx(x)
                    const double ar = c.alpha * r;
                    if (CK == CK_WOLF) {
                        const double ex = exp(-ar * ar);
                        const double erfcv = ex * erfcx_tab(ecx, ar, a.ecx_inv_h);
                        e = qiqj * (erfcv * inv_r - c.e_const) / kFourPiEps0;
                        const double factor = erfcv * inv_r * inv_r + c.alpha * kFrac2SqrtPi * ex * inv_r;
                        f = qiqj * (factor - c.f_const) * inv_r / kFourPiEps0;
                    } else {
                        const double ex = exp(-c.alpha * c.alpha * r * r);
                        const double erfcv = ex * erfcx_tab(ecx, ar, a.ecx_inv_h);
                        const double g = c.alpha * kFrac2SqrtPi * ex;
                        const double pref = qiqj / kFourPiEps0;
                        e = pref * erfcv * inv_r;
                        f = pref * inv_r * inv_r * (g + erfcv * inv_r);
                    }
                } else if (CK == CK_WOLF) {
                    wolf_pair(c, qiqj, r, inv_r, e, f);
                } else {
                    ewald_real(c.alpha, qiqj, r, inv_r, CK == CK_EWALD_ERF, e, f);
                }
                fs += f;
                e_c = e;
            }
            fx += fs * dx;
            fy += fs * dy;
            fz += fs * dz;
            if (kEnergy && j > i) {
                e_pair += e_p;
                e_coul += e_c;
                wxx += fs * dx * dx;
                wyy += fs * dy * dy;
                wzz += fs * dz * dz;
                wxy += fs * dx * dy;
                wxz += fs * dx * dz;
                wyz += fs * dy * dz;
            }
        };
        int k = sub;
        int eA = jat(k), eB = jat(k + TPA);  // list entries: atom and kind
        double4 pA = a.pos[nbr_atom(eA)], pB = a.pos[nbr_atom(eB)];
        for (; k < nn; k += 2 * TPA) {
            const int eA2 = jat(k + 2 * TPA), eB2 = jat(k + 3 * TPA);
            pair(nbr_atom(eA), pA, nbr_kind(eA));
            pA = a.pos[nbr_atom(eA2)];
            if (k + TPA < nn) pair(nbr_atom(eB), pB, nbr_kind(eB));
            pB = a.pos[nbr_atom(eB2)];
            eA = eA2;
            eB = eB2;
        }
    }
#pragma unroll
    for (int off = TPA / 2; off > 0; off >>= 1) {
        fx += __shfl_xor(fx, off, 64);
        fy += __shfl_xor(fy, off, 64);
        fz += __shfl_xor(fz, off, 64);
    }
    if (t < a.n && sub == 0) a.force[i] = make_double4(fx, fy, fz, 0.0);
    if (kEnergy) {
        double v[8] = {e_pair, e_coul, wxx, wyy, wzz, wxy, wxz, wyz};
        int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            double sum = wave_sum(v[q]);
            if (lane == 0) red[wid][q] = sum;
        }
        __syncthreads();
        if (threadIdx.x < 8) {
            double sum = 0;
            for (int w = 0; w < kPairBlock / 64; w++) sum += red[w][threadIdx.x];
            a.epart[blockIdx.x * 8 + threadIdx.x] = sum;
        }
    }
}

// Mixed form of k_pairs: positions are read as 32-bit fixed-point
// fractions (frac_pos, 16-byte gathers) whose wrapped difference is the
// minimum image, rounded to FP32 once; distance, LJ, erfc and the
// Ewald/Wolf force factor are FP32; the FP32 pair forces of each pass (two
// pairs) are added in FP32 and accumulated in FP64 (two operations per
// component and pass, against seven per pair for a compensated FP32 sum).  Energies and virial (energy mode) are
// accumulated in FP64.
//
// The kernel issues few instructions per pair (it is bound by the vector
// ALU when several systems share the device): a kind pair is one 16-byte
// LDS record {48 eps s^12, 24 eps s^6, rc^2 (0 when the kind pair has no
// term: Null, or excluded between molecules), shift}, the LJ force is
// (A r^-6 - B) r^-8 evaluated for every pair and selected by r^2 < rc^2, so
// that a Null or excluded kind pair adds an exact 0; Harmonic pair terms,
// rare, sit in a second record read only when the system has one (a
// uniform flag).  The Coulomb term is evaluated for every pair and selected
// by r <= rc_coulomb (q_i q_j = 0 gives an exact 0); in the Ewald and Wolf
// instances its r-dependence is h(r) r^-3 with h(r) = erfc(alpha r) + (2
// alpha / sqrt(pi)) r exp(-alpha^2 r^2) read from a table of FP32 pieces
// (fitted on the host to 1e-7 relative), one lookup and no exponential;
// energies keep exp(-x^2) erfcx(x).  1/r is the hardware reciprocal square
// root (1 ulp).
//
// The launch carries the rho(k) reduction blocks but not the
// intra-molecular terms (k_intra_terms, a launch of its own): their code
// needs 104 VGPRs, which would hold every pair block of the launch to four
// waves per SIMD, against eight for the pair code alone.
template <bool kEnergy, int TPA, int CK>
__global__ __launch_bounds__(kPairBlock) void k_pairs_mixed(PairArgs a) {
    __shared__ float4 ljt[kMaxKinds * kMaxKinds];
    __shared__ float2 hmt[kMaxKinds * kMaxKinds];  // Harmonic {k, x0}, k = 0 otherwise
    __shared__ double red[kPairBlock / 64][8];
    constexpr bool kTab = CK == CK_EWALD || CK == CK_WOLF;
    // erfcx pieces (energies, and forces without the h table) and h pieces
    __shared__ __align__(16) float ecx[kTab ? kErfcxPieces * (kErfcxDeg32 + 1) : 4];
    __shared__ __align__(16) float hft[kTab ? kErfcxPieces * (kErfcxDeg32 + 1) : 4];
    if ((int)blockIdx.x >= a.pair_blocks) {  // uniform over the block
        rho_reduce_block<kEnergy>(a.rr, blockIdx.x - a.pair_blocks);
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.rebuild) *a.rebuild = 0;
    const int nk2 = a.nkinds * a.nkinds;
    int harmonic = 0;
    for (int x = threadIdx.x; x < nk2; x += blockDim.x) {
        const PairParam pp = a.table[x];
        const bool live = pp.pot > POT_NULL && !pp.inter_excl;
        const float rc = (float)pp.rc;
        const bool lj = pp.pot == POT_LJ;
        const double s6 = lj ? pp.p0 * pp.p0 * pp.p0 * pp.p0 * pp.p0 * pp.p0 : 0.0;
        ljt[x] = make_float4(lj ? (float)(48.0 * pp.p1 * s6 * s6) : 0.f, lj ? (float)(24.0 * pp.p1 * s6) : 0.f,
                             live ? rc * rc : 0.f, (float)pp.shift);
        const bool hm = live && pp.pot == POT_HARMONIC;
        hmt[x] = make_float2(hm ? (float)pp.p0 : 0.f, hm ? (float)pp.p1 : 0.f);
        harmonic |= hm;
    }
    // uniform: the tables are there (the h table implies the erfcx table)
    const bool use_h = kTab && a.hf32 != nullptr && a.ecx32 != nullptr;
    const bool use_tab = kTab && a.ecx32 != nullptr;
    static_assert(kErfcxDeg32 + 1 == 4, "a piece of the FP32 tables is one float4");
    if (use_tab && (kEnergy || !use_h)) {
        // one piece (4 coefficients) per 16-byte copy
        for (int x = threadIdx.x; x < kErfcxPieces; x += blockDim.x)
            reinterpret_cast<float4*>(ecx)[x] = reinterpret_cast<const float4*>(a.ecx32)[x];
    }
    if (use_h) {
        for (int x = threadIdx.x; x < kErfcxPieces; x += blockDim.x)
            reinterpret_cast<float4*>(hft)[x] = reinterpret_cast<const float4*>(a.hf32)[x];
    }
    const float hf_inv_h = (float)a.hf_inv_h;
    harmonic = __syncthreads_or(harmonic);  // also the barrier of the LDS tables
    const float ecx_inv_h = (float)a.ecx_inv_h;
    const float alpha = (float)a.coul.alpha;
    const float rcc2 = (float)(a.coul.rc * a.coul.rc);
    const float g0 = (float)(a.coul.alpha * kFrac2SqrtPi);
    const float inv4pe = (float)(1.0 / kFourPiEps0);
    const float we = (float)a.coul.e_const, wf = (float)a.coul.f_const;
    const bool periodic = a.box.periodic;  // uniform
    // L / 2^32: one unit of the fixed-point fractions
    const float Lx = (float)(a.box.L[0] * 0x1p-32), Ly = (float)(a.box.L[1] * 0x1p-32),
                Lz = (float)(a.box.L[2] * 0x1p-32);
    const float4* __restrict__ frac = a.frac;

    // XCD-contiguous order of the pair blocks (the launch has a multiple of 8)
    int g = xcd_contiguous(blockIdx.x, a.pair_blocks) * blockDim.x + threadIdx.x;
    int t = g / TPA;
    int sub = g % TPA;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    double e_pair = 0, e_coul = 0;
    double wxx = 0, wyy = 0, wzz = 0, wxy = 0, wxz = 0, wyz = 0;
    int i = -1;
    if (t < a.n) {
        i = a.sorted[t];
        const float4 pi = frac[i];
        const int ki = a.meta[i].x;
        const float qi = pi.w;
        // the home charge with 1 / (4 pi eps0): one product per pair fewer
        const float qi_k = qi * inv4pe;
        const int nn = a.nnbr[t];
        const int* row = a.nbr + (size_t)t * a.cap;
        const float4* __restrict__ ljrow = ljt + ki * a.nkinds;
        const float2* __restrict__ hmrow = hmt + ki * a.nkinds;
        // Software pipeline over the row: the row entries two passes ahead
        // and the next pass's neighbour positions are loaded while the
        // current pair is evaluated, unrolled by two on alternating
        // registers (A, B); an entry carries the neighbour's kind.  Entries
        // past the row end read a clamped index and stand for atom i
        // itself, never evaluated.
        auto jat = [&](int k) {
            int v = row[min(k, max(nn - 1, 0))];
            return k < nn ? v : nbr_entry(i, ki);
        };
        // one list entry: returns the force factor f/r, with the difference
        // vector in dx, dy, dz (the energies and virial, in energy mode,
        // are accumulated here in FP64)
        auto pair = [&](int j, const float4& pj, int kj, float& dx, float& dy, float& dz) -> float {
            if (periodic) {
                // the wrapped 32-bit difference is the minimum image
                dx = (float)(int)(__float_as_uint(pi.x) - __float_as_uint(pj.x)) * Lx;
                dy = (float)(int)(__float_as_uint(pi.y) - __float_as_uint(pj.y)) * Ly;
                dz = (float)(int)(__float_as_uint(pi.z) - __float_as_uint(pj.z)) * Lz;
            } else {
                dx = pi.x - pj.x;
                dy = pi.y - pj.y;
                dz = pi.z - pj.z;
            }
            const float r2 = dx * dx + dy * dy + dz * dz;
            const float inv_r = __builtin_amdgcn_rsqf(r2);
            const float inv_r2 = inv_r * inv_r;
            const float4 P = ljrow[kj];
            const bool in_pair = r2 < P.z;
            const float i6 = inv_r2 * inv_r2 * inv_r2;
            float fp = (P.x * i6 - P.y) * (i6 * inv_r2);
            float e_p = 0.f, e_c = 0.f;
            if (kEnergy) e_p = i6 * (P.x * (1.f / 12.f) * i6 - P.y * (1.f / 6.f));
            if (harmonic) {  // uniform
                const float2 H = hmrow[kj];
                const float r = r2 * inv_r;
                fp += H.x * (H.y - r) * inv_r;
                if (kEnergy) e_p += 0.5f * H.x * (r - H.y) * (r - H.y);
            }
            float fs = in_pair ? fp : 0.f;
            if (kEnergy) e_p = in_pair ? e_p - P.w : 0.f;
            const float pref = qi_k * pj.w;  // q_i q_j / (4 pi eps0)
            if (kTab && use_h) {
                // forces from the h table: no exponential, one lookup
                const float r = r2 * inv_r;
                const float h = erfcx_tab(hft, r, hf_inv_h);
                const float inv_r3 = inv_r * inv_r2;
                const float fc = CK == CK_EWALD ? pref * (h * inv_r3) : pref * (h * inv_r3 - wf * inv_r);
                const bool in_coul = r2 <= rcc2;
                fs += in_coul ? fc : 0.f;
                if (kEnergy) {
                    const float x = alpha * r;
                    const float erfcv = __expf(-x * x) * erfcx_tab(ecx, x, ecx_inv_h);
                    e_c = in_coul ? (CK == CK_EWALD ? pref * erfcv * inv_r : pref * (erfcv * inv_r - we)) : 0.f;
                }
            } else if (CK != CK_NONE) {
                const float r = r2 * inv_r;
                float ex2, erfcv;
                if (kTab && use_tab) {
                    const float x = alpha * r;
                    ex2 = __expf(-x * x);
                    erfcv = ex2 * erfcx_tab(ecx, x, ecx_inv_h);
                } else {
                    erfcv = erfc_cheb(alpha * r, ex2);
                }
                float fc;
                if (CK == CK_EWALD_ERF) {
                    const float erfv = 1.f - erfcv;
                    fc = pref * inv_r2 * (g0 * ex2 - erfv * inv_r);
                    if (kEnergy) e_c = -pref * erfv * inv_r;
                } else if (CK == CK_EWALD) {
                    fc = pref * inv_r2 * (g0 * ex2 + erfcv * inv_r);
                    if (kEnergy) e_c = pref * erfcv * inv_r;
                } else {
                    const float factor = erfcv * inv_r2 + g0 * ex2 * inv_r;
                    fc = pref * (factor - wf) * inv_r;
                    if (kEnergy) e_c = pref * (erfcv * inv_r - we);
                }
                const bool in_coul = r2 <= rcc2;
                fs += in_coul ? fc : 0.f;
                if (kEnergy) e_c = in_coul ? e_c : 0.f;
            }
            if (kEnergy && j > i) {
                e_pair += e_p;
                e_coul += e_c;
                const double f = fs, ddx = dx, ddy = dy, ddz = dz;
                wxx += f * ddx * ddx;
                wyy += f * ddy * ddy;
                wzz += f * ddz * ddz;
                wxy += f * ddx * ddy;
                wxz += f * ddx * ddz;
                wyz += f * ddy * ddz;
            }
            return fs;
        };
        int k = sub;
        int eA = jat(k), eB = jat(k + TPA);  // list entries: atom and kind
        float4 pA = frac[nbr_atom(eA)], pB = frac[nbr_atom(eB)];
        for (; k < nn; k += 2 * TPA) {
            const int eA2 = jat(k + 2 * TPA), eB2 = jat(k + 3 * TPA);
            // the two pair forces of a pass are added in FP32 (one rounding
            // of their sum), then to the FP64 accumulators
            float ax, ay, az;
            const float fa = pair(nbr_atom(eA), pA, nbr_kind(eA), ax, ay, az);
            float tx = fa * ax, ty = fa * ay, tz = fa * az;
            pA = frac[nbr_atom(eA2)];
            if (k + TPA < nn) {
                float bx, by, bz;
                const float fb = pair(nbr_atom(eB), pB, nbr_kind(eB), bx, by, bz);
                tx += fb * bx;
                ty += fb * by;
                tz += fb * bz;
            }
            pB = frac[nbr_atom(eB2)];
            fx += (double)tx;
            fy += (double)ty;
            fz += (double)tz;
            eA = eA2;
            eB = eB2;
        }
    }
    double sx = fx, sy = fy, sz = fz;
#pragma unroll
    for (int off = TPA / 2; off > 0; off >>= 1) {
        sx += __shfl_xor(sx, off, 64);
        sy += __shfl_xor(sy, off, 64);
        sz += __shfl_xor(sz, off, 64);
    }
    if (t < a.n && sub == 0) a.force[i] = make_double4(sx, sy, sz, 0.0);
    if (kEnergy) {
        double v[8] = {e_pair, e_coul, wxx, wyy, wzz, wxy, wxz, wyz};
        int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            double sum = wave_sum(v[q]);
            if (lane == 0) red[wid][q] = sum;
        }
        __syncthreads();
        if (threadIdx.x < 8) {
            double sum = 0;
            for (int w = 0; w < kPairBlock / 64; w++) sum += red[w][threadIdx.x];
            a.epart[blockIdx.x * 8 + threadIdx.x] = sum;
        }
    }
}

// Pair-set export: (i < j) with r < rc from the neighbour lists, using the
// same distance code as k_pairs.
__global__ void k_pairs_within(const double4* __restrict__ pos, const int* __restrict__ sorted,
                               const int* __restrict__ nbr, const int* __restrict__ nnbr, int n, int cap, Box box,
                               double rc,
                               uint2* __restrict__ out, unsigned long long out_cap,
                               unsigned long long* __restrict__ count) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i = sorted[t];
    double4 pi = pos[i];
    int nn = nnbr[t];
    for (int k = 0; k < nn; k++) {
        int j = nbr_atom(nbr[(size_t)t * cap + k]);
        if (j <= i) continue;
        double4 pj = pos[j];
        double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        min_image(box, dx, dy, dz);
        double r = sqrt(norm2_ordered(dx, dy, dz));
        if (r < rc) {
            unsigned long long idx = atomicAdd(count, 1ull);
            if (idx < out_cap) out[idx] = make_uint2((unsigned)i, (unsigned)j);
        }
    }
}


// Total force on atom i: the non-bonded pair force, the intra-molecular slot
// contributions in slot order, and the k-space partials
// q_i / (4 pi eps0) sum_g fk[g][i] (ewald.rs:716-719).  Shared by the MD
// second half and k_force_combine so both paths add in the same order.
__device__ __forceinline__ double4 total_force(int i, int n, const double4* __restrict__ force,
                                               const double4* __restrict__ contrib, const int* __restrict__ slot_start,
                                               const double4* __restrict__ fk, int ksplit,
                                               const double4* __restrict__ pos) {
    double4 f = force[i];
    // loops unrolled so that their loads are in flight together (a split
    // at a time cost one L2 round trip per split); the additions keep
    // their order
    if (slot_start) {
#pragma unroll 4
        for (int sl = slot_start[i]; sl < slot_start[i + 1]; sl++) {
            double4 c = contrib[sl];
            f.x += c.x;
            f.y += c.y;
            f.z += c.z;
        }
    }
    if (ksplit > 0) {
        double sx = 0, sy = 0, sz = 0;
#pragma unroll 8
        for (int g = 0; g < ksplit; g++) {
            double4 p = fk[(size_t)g * n + i];
            sx += p.x;
            sy += p.y;
            sz += p.z;
        }
        double q = pos[i].w;
        f.x += (q * sx) / kFourPiEps0;
        f.y += (q * sy) / kFourPiEps0;
        f.z += (q * sz) / kFourPiEps0;
    }
    return f;
}

__global__ void k_force_combine(double4* __restrict__ force, const double4* __restrict__ contrib,
                                const int* __restrict__ slot_start, const double4* __restrict__ fk, int ksplit,
                                const double4* __restrict__ pos, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    force[i] = total_force(i, n, force, contrib, slot_start, fk, ksplit, pos);
}

// ---------------------------------------------------------------------------------
// MolecularVirial (compute.rs:278-364).  k_mol_com: centre of mass of each
// molecule, masses and positions summed in atom order (molecules.rs:256-263).
// k_mol_pairs: every inter-molecular pair of the Verlet list once (the atom
// of the lower-numbered molecule owns it) adds w_ab (r_ab . r_ij) / r_ab^2,
// w_ab being the pair virial of the pair potential (compute.rs:289-307), the
// Ewald real space (ewald.rs:533-570) or Wolf (wolf.rs:285-324) and r_ij the
// minimum image of the two centres of mass.  k_mol_kspace: the k-space
// correction sum_i F_i (x) (r_i - com) of ewald.rs:736-753.  Both write
// per-block partial sums that the host adds in block order.
// ---------------------------------------------------------------------------------
__global__ void k_mol_com(const double4* __restrict__ pos, const double4* __restrict__ vel,
                          const int* __restrict__ mol_start, int nmol, double4* __restrict__ com) {
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    double total = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (int i = mol_start[m]; i < mol_start[m + 1]; i++) {
        const double4 p = pos[i];
        const double mass = vel[i].w;
        total += mass;
        cx += mass * p.x;
        cy += mass * p.y;
        cz += mass * p.z;
    }
    com[m] = make_double4(cx / total, cy / total, cz / total, total);
}

struct MolVirialArgs {
    const double4* pos;
    const int2* meta;
    const int* sorted;
    const int* nbr;
    const int* nnbr;
    int cap, n, nkinds;
    const PairParam* table;
    const double* tabs;
    CoulParam coul;
    Box box;
    const double4* com;
    double* part;  // 6 per block: xx, yy, zz, xy, xz, yz
};

__device__ __forceinline__ void block_sum_store(double* v, int nv, double* out) {
    __shared__ double red[4][9];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int q = 0; q < nv; q++) {
        double sum = wave_sum(v[q]);
        if (lane == 0) red[wid][q] = sum;
    }
    __syncthreads();
    if (threadIdx.x < nv) {
        double sum = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) sum += red[w][threadIdx.x];
        out[blockIdx.x * nv + threadIdx.x] = sum;
    }
}

__global__ __launch_bounds__(256) void k_mol_pairs(MolVirialArgs a) {
    double w[6] = {0, 0, 0, 0, 0, 0};
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.n) {
        const int i = a.sorted[t];
        const double4 pi = a.pos[i];
        const int2 mi = a.meta[i];
        const double4 ci = a.com[mi.y];
        const int* row = a.nbr + (size_t)t * a.cap;
        const int nn = a.nnbr[t];
        for (int k = 0; k < nn; k++) {
            const int e = row[k];
            const int j = nbr_atom(e);
            const int mj = a.meta[j].y;
            if (mj <= mi.y) continue;
            const double4 pj = a.pos[j];
            double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            min_image(a.box, dx, dy, dz);
            const double r2 = norm2_ordered(dx, dy, dz);
            const double r = sqrt(r2);
            const double inv_r = 1.0 / r;
            double fs = 0.0;
            const PairParam& pp = a.table[mi.x * a.nkinds + nbr_kind(e)];
            if (pp.pot > POT_NULL && !pp.inter_excl && r < pp.rc) {
                double en, f;
                pot_eval_any(pp, a.tabs, r, inv_r, en, f);
                fs += f;
            }
            const double qq = pi.w * pj.w;
            if (a.coul.kind != COUL_NONE && qq != 0.0 && r <= a.coul.rc) {
                double en, f;
                if (a.coul.kind == COUL_EWALD) {
                    ewald_real(a.coul.alpha, qq, r, inv_r, a.coul.inter_excl != 0, en, f);
                    fs += f;
                } else if (!a.coul.inter_excl) {
                    wolf_pair(a.coul, qq, r, inv_r, en, f);
                    fs += f;
                }
            }
            if (fs == 0.0) continue;
            const double4 cj = a.com[mj];
            double rx = ci.x - cj.x, ry = ci.y - cj.y, rz = ci.z - cj.z;
            min_image(a.box, rx, ry, rz);
            const double wgt = (dx * rx + dy * ry + dz * rz) / r2;
            w[0] += fs * dx * dx * wgt;
            w[1] += fs * dy * dy * wgt;
            w[2] += fs * dz * dz * wgt;
            w[3] += fs * dx * dy * wgt;
            w[4] += fs * dx * dz * wgt;
            w[5] += fs * dy * dz * wgt;
        }
    }
    block_sum_store(w, 6, a.part);
}

__global__ __launch_bounds__(256) void k_mol_kspace(const double4* __restrict__ pos, const int2* __restrict__ meta,
                                                     const double4* __restrict__ fk, int ksplit, int n,
                                                     const double4* __restrict__ com, double* __restrict__ part) {
    double c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double sx = 0, sy = 0, sz = 0;
        for (int g = 0; g < ksplit; g++) {
            const double4 p = fk[(size_t)g * n + i];
            sx += p.x;
            sy += p.y;
            sz += p.z;
        }
        const double4 p = pos[i];
        const double fx = (p.w * sx) / kFourPiEps0, fy = (p.w * sy) / kFourPiEps0, fz = (p.w * sz) / kFourPiEps0;
        const double4 cm = com[meta[i].y];
        const double d[3] = {p.x - cm.x, p.y - cm.y, p.z - cm.z};
        const double f[3] = {fx, fy, fz};
        for (int u = 0; u < 3; u++)
            for (int v = 0; v < 3; v++) c[3 * u + v] = f[u] * d[v];
    }
    block_sum_store(c, 9, part);
}

// ---------------------------------------------------------------------------------
// Ewald k-space (ewald.rs:633-733).  k_phase computes, once per evaluation,
// e^{i n k_axis . r_i} for every atom and axis by the reference's recursion
// (polar for n = 1, products for |n| >= 2, conjugates for n < 0) into a
// table in HBM; the rho and force kernels stage the rows they use in LDS and
// form e^{i k.r} = (Ex * Ey) * Ez as ewald.rs:666-668.
// ---------------------------------------------------------------------------------

// Phase-table rows: axis 0 keeps n = 0..kmax (ikx >= 0 in the half-space
// set), axes 1 and 2 keep n = -kmax..kmax: 5*kmax + 3 rows of `ld` atoms.
__host__ __device__ __forceinline__ int phase_row(int axis, int n, int kmax) {
    return axis == 0 ? n : (kmax + 1) + (axis - 1) * (2 * kmax + 1) + (n + kmax);
}
__host__ __device__ __forceinline__ int phase_rows(int kmax) { return 5 * kmax + 3; }

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// XCD-aware coordinates of a logical 2-D grid (gx, gy) launched as a 1-D
// grid of gx * round_up(gy, 8) blocks.  The dispatcher hands block L to XCD
// L mod 8, and each XCD has its own L2; logical row `by` is given only ids
// with L mod 8 == by mod 8, so all blocks that read row by's data (an atom
// chunk of the phase table) share one L2 and the chunk is fetched from HBM
// once.  False for the padding blocks (whole block; call before any barrier).
__device__ __forceinline__ bool xcd_block_at(int L, int gx, int gy, int& bx, int& by) {
    int q = L >> 3;
    by = (q / gx) * 8 + (L & 7);
    bx = q - (q / gx) * gx;
    return by < gy;
}
__device__ __forceinline__ bool xcd_block(int gx, int gy, int& bx, int& by) {
    return xcd_block_at(blockIdx.x, gx, gy, bx, by);
}
__host__ __forceinline__ unsigned xcd_grid(int gx, int gy) { return unsigned(gx) * unsigned((gy + 7) / 8 * 8); }

// Rows of atom column `a` along `axis`: e^{i n k_axis x} by the reference's
// recursion; k_vector(e_axis) . r = (2 pi * inv_axis) * x  (cells.rs:224).
__device__ __forceinline__ void phase_axis64(double x, double inv, int axis, int kmax, double2* __restrict__ tab,
                                             int ld, int a) {
    double phase = (2.0 * kPi * inv) * x;
    double sn, cs;
    sincos(phase, &sn, &cs);
    double2 e1 = make_double2(cs, sn);
    tab[(size_t)phase_row(axis, 0, kmax) * ld + a] = make_double2(1.0, 0.0);
    tab[(size_t)phase_row(axis, 1, kmax) * ld + a] = e1;
    if (axis != 0) tab[(size_t)phase_row(axis, -1, kmax) * ld + a] = make_double2(e1.x, -e1.y);
    double2 e = e1;
    for (int k = 2; k <= kmax; k++) {
        e = make_double2(e.x * e1.x - e.y * e1.y, e.x * e1.y + e.y * e1.x);
        tab[(size_t)phase_row(axis, k, kmax) * ld + a] = e;
        if (axis != 0) tab[(size_t)phase_row(axis, -k, kmax) * ld + a] = make_double2(e.x, -e.y);
    }
}

// Mixed form: each phase n k_axis x formed and reduced to [-pi, pi] in FP64,
// then sincosf.
__device__ __forceinline__ float2 phase32_n(double phase, int k) {
    const double two_pi = 2.0 * kPi, inv_two_pi = 1.0 / (2.0 * kPi);
    double th = k * phase;
    th -= two_pi * rint(th * inv_two_pi);
    float sn, cs;
    sincosf((float)th, &sn, &cs);
    return make_float2(cs, sn);
}

__device__ __forceinline__ void phase_axis32(double x, double inv, int axis, int kmax, float2* __restrict__ tab,
                                             int ld, int a, int k0 = 1, int k1 = -1) {
    double phase = (2.0 * kPi * inv) * x;
    if (k1 < 0) k1 = kmax;
    if (k0 == 1) tab[(size_t)phase_row(axis, 0, kmax) * ld + a] = make_float2(1.f, 0.f);
    for (int k = k0; k <= k1; k++) {
        const float2 e = phase32_n(phase, k);
        tab[(size_t)phase_row(axis, k, kmax) * ld + a] = e;
        if (axis != 0) tab[(size_t)phase_row(axis, -k, kmax) * ld + a] = make_float2(e.x, -e.y);
    }
}

// Where the MD step's first half writes the phase table of the new
// positions (mode 0: no k-space, 1: FP64 table, 2: FP32 table).
struct PhaseOut {
    int mode, kmax, ld;
    Box box;
    double2* tab;
    double* q;
    float2* tab32;
    float* q32;
};

__device__ __forceinline__ void phase_atom(const PhaseOut& ph, int a, bool live, double4 p) {
    double xyz[3] = {live ? p.x : 0.0, live ? p.y : 0.0, live ? p.z : 0.0};
    if (ph.mode == 1) {
        for (int axis = 0; axis < 3; axis++) phase_axis64(xyz[axis], ph.box.inv[axis], axis, ph.kmax, ph.tab, ph.ld, a);
        ph.q[a] = live ? p.w : 0.0;
    } else {
        for (int axis = 0; axis < 3; axis++)
            phase_axis32(xyz[axis], ph.box.inv[axis], axis, ph.kmax, ph.tab32, ph.ld, a);
        ph.q32[a] = live ? (float)p.w : 0.f;
    }
}

// One thread per (atom, axis); tab[row * ld + atom], q[atom] (0 past n).
__global__ void k_phase(const double4* __restrict__ pos, int n, int ld, Box b, int kmax, double2* __restrict__ tab,
                        double* __restrict__ q) {
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    int axis = gid / ld;
    int a = gid - axis * ld;
    if (axis >= 3) return;
    bool live = a < n;
    double x = 0.0, qa = 0.0;
    if (live) {
        double4 p = pos[a];
        x = axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
        qa = p.w;
    }
    phase_axis64(x, b.inv[axis], axis, kmax, tab, ld, a);
    if (axis == 0) q[a] = live ? qa : 0.0;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// rho partial sums by row units.  A unit is up to kRhoUnitK (5) consecutive
// k-vectors of one row (fixed ikx, iky; consecutive ikz): per atom, the lane
// forms q e^{i k0.r} = q (Ex * Ey) * Ez[ikz0] and steps to the next ikz by
// one product with Ez[1], 7.2 operations and 0.6 table reads per (atom, k).
// Five per unit pads the k set least (units fill whole 64-lane blocks more
// often than with four: SPC/E kmax 10, 480 units in 512 lanes against 589
// in 640; 8.9 against 9.2 operations per k-vector and atom, lane padding
// included).  Block = 64 units (one per lane) x 8 waves; the block's atom
// chunk of `width` atoms is split in eight and each wave sums its eighth;
// the eight wave sums are added in order in LDS and wave j < 5 writes
// k-vector j of every unit: partial[chunk * 5 * nunits + j * nunits + unit]
// (slots with j past the unit's count are not written).  Table rows are
// read straight from L2: a wave's lanes read few distinct rows at one atom
// column.  The table is padded with zero charges to `ld` atoms.  Grid:
// xcd_grid(nub, nchunks), so all unit blocks of an atom chunk share one L2.
// Mixed form (T = float): each wave's eighth is summed in FP32, the rest in
// FP64.
constexpr int kRhoUnitK = 5;
struct KUnit {
    short ikx, iky, kz0, nv;  // row, first ikz, k-vectors in the unit (1..kRhoUnitK)
    int kb, pad;              // index of the unit's first k-vector
};

template <typename T, typename T2>
struct RhoRowsArgs {
    const T2* tab;
    const T* q;
    int ld, kmax, width;
    const KUnit* units;
    int nunits;
    double2* partial;
    int nub, nchunks;
    int stage;  // 1: phase rows staged in LDS when they fit (rho_kind 1), 0: read from the table (rho_kind 3)
};

// Phase rows the staged row kernel loads per round; the phase tables are
// allocated with at least this many rows (kmax <= 12 staged).
constexpr int kRhoStageRows = 64;
// Atoms per staging round of the row kernel: every phase row of S
// consecutive atoms is copied to LDS at once.
template <typename T>
constexpr int rho_stage_atoms() { return sizeof(T) == 4 ? 8 : 4; }
// LDS bytes of one wave's staging area (rows padded to S + 1 entries so that
// lanes reading different rows at one atom fall in different banks; the
// charges follow the rows), rounded to 16 bytes.
template <typename T, typename T2>
__host__ __device__ constexpr size_t rho_stage_wave_bytes(int kmax) {
    return ((size_t)phase_rows(kmax) * (rho_stage_atoms<T>() + 1) * sizeof(T2) + rho_stage_atoms<T>() * sizeof(T) +
            15) / 16 * 16;
}

// NW waves per block (8 in k_rho_rows, k_cells_rho and k_direct_rho),
// logical block index L of the xcd_grid(nub, nchunks) layout.  buf: the
// launch's LDS (buf_bytes), which holds the eight wave sums at the end and,
// when the staging areas of all NW waves fit in it, the staged phase rows.
// kPrefetch: the staged loop loads the next round while it sums the current
// one (more registers held: used where the launch's LDS, not its registers,
// sets the residency).
template <typename T, typename T2, int NW, bool kPrefetch = false>
__device__ __forceinline__ void rho_rows_block(const RhoRowsArgs<T, T2>& ra, int L, unsigned char* buf,
                                               size_t buf_bytes) {
    static_assert(NW >= kRhoUnitK, "one wave writes each k-vector of a unit");
    double2(*red)[kRhoUnitK][64] = reinterpret_cast<double2(*)[kRhoUnitK][64]>(buf);
    const T2* __restrict__ tab = ra.tab;
    const T* __restrict__ q = ra.q;
    const int ld = ra.ld, kmax = ra.kmax, width = ra.width, nunits = ra.nunits;
    const KUnit* __restrict__ units = ra.units;
    double2* __restrict__ partial = ra.partial;
    int ub, chunk;
    if (!xcd_block_at(L, ra.nub, ra.nchunks, ub, chunk)) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int u = ub * 64 + lane;
    const KUnit un = units[u < nunits ? u : nunits - 1];
    const int per = width / NW;
    // the wave's atom range, visibly uniform.  a1 - a0 is a multiple of 8
    // or <= 0: widths are multiples of 64 (as is ld), so per and a0 are
    // multiples of 8
    const int a0 = __builtin_amdgcn_readfirstlane(chunk * width + w * per);
    const int a1 = min(a0 + per, ld);
    T re[kRhoUnitK], im[kRhoUnitK];
#pragma unroll
    for (int v = 0; v < kRhoUnitK; v++) re[v] = im[v] = 0;
    // one atom's terms, in the order of the reference's triple product
    auto add_atom = [&](T2 x, T2 y, T2 z, T2 e1, T qa) {
        T2 e = cmul(cmul(x, y), z);
        e.x *= qa;
        e.y *= qa;
#pragma unroll
        for (int v = 0; v < kRhoUnitK; v++) {
            if (v > 0) e = cmul(e, e1);
            re[v] += e.x;
            im[v] += e.y;
        }
    };
    constexpr int S = rho_stage_atoms<T>(), SP = S + 1;
    constexpr int kStep = 64 / S;  // rows per load instruction of a wave
    const size_t wave_bytes = rho_stage_wave_bytes<T, T2>(kmax);
    const int prow = phase_rows(kmax);
    if (ra.stage && (size_t)NW * wave_bytes <= buf_bytes && prow <= kRhoStageRows) {
        // Staged: each round copies every phase row of the next S atoms
        // (and their charges) to the wave's own LDS area, one load per 64
        // entries (S consecutive atoms of 64 / S rows), all in flight
        // together; then each lane reads its unit's rows from LDS: one
        // memory latency per S atoms.  Atoms are added in index order, as in
        // the unstaged loop.  The loads cover kRhoStageRows rows, which the
        // table is allocated to hold; rows past the phase rows are not
        // stored.  Only this wave reads its area, and a wave's LDS
        // operations complete in order, so a round's reads precede the next
        // round's writes.
        T2* tile = reinterpret_cast<T2*>(buf + (size_t)w * wave_bytes);
        T* qt = reinterpret_cast<T*>(tile + (size_t)prow * SP);
        const int rx = phase_row(0, un.ikx, kmax) * SP, ry = phase_row(1, un.iky, kmax) * SP;
        const int rz = phase_row(2, un.kz0, kmax) * SP, r1 = phase_row(2, 1, kmax) * SP;
        const int lr = lane / S, lj = lane - lr * S, lq = min(lane, S - 1);
        const T2* src = tab + (size_t)lr * ld + lj;
        constexpr int kR = kRhoStageRows / kStep;
        T2 v[kR];
        T qv;
        auto load = [&](int a) {
#pragma unroll
            for (int it = 0; it < kR; it++) v[it] = src[(size_t)(it * kStep) * ld + a];
            qv = q[a + lq];
        };
        auto stage = [&]() {
#pragma unroll
            for (int it = 0; it < kR; it++) {
                const int r = lr + it * kStep;
                if (r < prow) tile[r * SP + lj] = v[it];
            }
            qt[lq] = qv;
            // the wave's LDS writes done before any lane reads them
            __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
            __builtin_amdgcn_wave_barrier();
        };
        auto sum = [&]() {
#pragma unroll
            for (int j = 0; j < S; j++) add_atom(tile[rx + j], tile[ry + j], tile[rz + j], tile[r1 + j], qt[j]);
            __builtin_amdgcn_wave_barrier();
        };
        if constexpr (kPrefetch) {
            // the next round's loads in flight while this round is summed
            // (one memory latency per wave).  The first round's values go
            // through an empty asm statement once they have landed, so that
            // the values entering the loop are not all loads and the
            // compiler keeps the early loads where they are; the last round
            // reloads its own atoms (in bounds, not used).
            if (a0 < a1) {
                load(a0);
#pragma unroll
                for (int it = 0; it < kR; it++) asm volatile("" : "+v"(v[it].x), "+v"(v[it].y));
                asm volatile("" : "+v"(qv));
            }
            for (int a = a0; a < a1; a += S) {
                stage();
                load(a + S < a1 ? a + S : a);
                sum();
            }
        } else {
            for (int a = a0; a < a1; a += S) {
                load(a);
                stage();
                sum();
            }
        }
    } else {
        const T2* cx = tab + (size_t)phase_row(0, un.ikx, kmax) * ld;
        const T2* cy = tab + (size_t)phase_row(1, un.iky, kmax) * ld;
        const T2* cz = tab + (size_t)phase_row(2, un.kz0, kmax) * ld;
        const T2* c1 = tab + (size_t)phase_row(2, 1, kmax) * ld;
#pragma unroll 4
        for (int a = a0; a < a1; a++) add_atom(cx[a], cy[a], cz[a], c1[a], q[a]);
    }
    // every wave is past its staging area before the sums overwrite it
    __syncthreads();
#pragma unroll
    for (int v = 0; v < kRhoUnitK; v++) red[w][v][lane] = make_double2(re[v], im[v]);
    __syncthreads();
    if (u < nunits && w < un.nv) {
        double2 sum = red[0][w][lane];
#pragma unroll
        for (int v = 1; v < NW; v++) {
            sum.x += red[v][w][lane].x;
            sum.y += red[v][w][lane].y;
        }
        partial[(size_t)chunk * kRhoUnitK * nunits + (size_t)w * nunits + u] = sum;
    }
}

// LDS of the eight wave sums of a row-kernel block
constexpr size_t kRhoRed = sizeof(double2) * (kCellBlock / 64) * kRhoUnitK * 64;

// Same 8-wave blocks as the rho part of k_cells_rho, so that the partial
// sums do not depend on which launch computed them.
template <typename T, typename T2>
__global__ __launch_bounds__(512) void k_rho_rows(RhoRowsArgs<T, T2> ra) {
    __shared__ __align__(16) unsigned char buf[kRhoRed];
    rho_rows_block<T, T2, 8>(ra, blockIdx.x, buf, sizeof(buf));
}

// One launch for two independent passes of a force evaluation: blocks
// [0, ncb) build the Verlet list (ncb = cells rounded up to 8, so that the
// rho blocks that follow keep the XCD layout of xcd_grid), the rest compute
// the rho(k) partials with 8 waves per block.  Both read only what the
// first half-step and the cell list wrote.
template <bool kMinImage, typename T, typename T2>
__global__ __launch_bounds__(kCellBlock) void k_cells_rho(CellsArgs ca, RhoRowsArgs<T, T2> ra, int ncb) {
    __shared__ __align__(16) unsigned char buf[sizeof(CellsLds) > kRhoRed ? sizeof(CellsLds) : kRhoRed];
    if ((int)blockIdx.x < ncb)
        nlist_cells_block<kMinImage>(ca, blockIdx.x, ncb, *reinterpret_cast<CellsLds*>(buf));
    else
        rho_rows_block<T, T2, kCellBlock / 64>(ra, blockIdx.x - ncb, buf, sizeof(buf));
}

// The same with the direct Verlet build in the first ncb blocks; the rho
// partials in the next nrho blocks, and any blocks past those run the
// intra-molecular terms (forces only, kCellBlock terms per block), which
// need only the positions of the first half-step.
// kIntra: the instance with the intra-molecular blocks (mixed form); the
// others leave their code out (it costs the rho blocks registers).
// Four waves per SIMD: its 64 KB of LDS allow two blocks (sixteen waves) per
// CU, so registers up to 128 cost nothing and above it would cost a block.
template <bool kMinImage, typename T, typename T2, bool kIntra = false>
__global__ __launch_bounds__(kCellBlock, 4) void k_direct_rho(DirectArgs da, RhoRowsArgs<T, T2> ra, int ncb, int nrho,
                                                           IntraArgs ia) {
    __shared__ __align__(16) unsigned char buf[sizeof(DirectLds) > kRhoRed ? sizeof(DirectLds) : kRhoRed];
    if ((int)blockIdx.x < ncb)
        nlist_direct_block<kMinImage>(da, blockIdx.x, ncb, *reinterpret_cast<DirectLds*>(buf));
    else if ((int)blockIdx.x < ncb + nrho)
        rho_rows_block<T, T2, kCellBlock / 64, true>(ra, blockIdx.x - ncb, buf, sizeof(buf));
    else if constexpr (kIntra)
        intra_block<false>(ia, blockIdx.x - ncb - nrho);
}


// ---- rho(k) on the matrix cores ---------------------------------------------------
// With e^{ik.r} = Ex[ikx] * Eyz[(iky, ikz)] and Eyz = q Ey Ez, the structure
// factor is a complex matrix product over atoms:
//     rho[ikx][col] = sum_i Ex[i][ikx] * Eyz[i][col],   col = (iky, ikz)
// (the same triple product as ewald.rs:664-673, grouped differently).  Rows
// ikx = 0..15 (kmax <= 16), columns = the distinct (iky, ikz) of the k set in
// tiles of 16; the reduction index runs over (atom, re/im), so one
// v_mfma 16x16x4 adds two atoms to a 16 x 16 tile of Re(rho) and another to
// Im(rho):
//     Re: A = (Ex.re, -Ex.im), B = (Eyz.re, Eyz.im)
//     Im: A = (Ex.im,  Ex.re), B = the same
// The reduction over atoms -- across lanes in a lane-per-atom layout --
// happens inside the matrix core.  Block = 8 waves on `width` atoms whose
// phase factors it forms from the positions into LDS as
// [atom][X 0..15 | Ey -kmax..kmax | Ez -kmax..kmax] plus the charges
// (applied to the A operand); wave
// (sub = w & 3, half = w >> 2) sums atoms [sub * width/4, +width/4) of each
// stage for 3 tiles; a block covers 6 tiles (group) and `cpb`
// consecutive stages (chunk).  The four sub-waves' tiles are added in order
// in FP64 and written to partial[chunk * nk + k] for the valid k only
// (kmap); the fixed-order chunk reduction of rho_reduce_block follows.
typedef float rho_f4 __attribute__((ext_vector_type(4)));
typedef double rho_d4 __attribute__((ext_vector_type(4)));
template <typename T> struct MfmaAcc;
template <> struct MfmaAcc<float> { typedef rho_f4 type; };
template <> struct MfmaAcc<double> { typedef rho_d4 type; };
__device__ __forceinline__ rho_f4 mfma16x16x4(float a, float b, rho_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ rho_d4 mfma16x16x4(double a, double b, rho_d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
// Both take A[i][k] from lane i + 16 k and B[k][j] from lane j + 16 k; output
// r of lane l is D[row][l % 16] with row = 4 (l / 16) + r for the FP32 form
// and l / 16 + 4 r for the FP64 form (measured on gfx950).
__device__ __forceinline__ int mfma16_row(float, int l16, int r) { return 4 * l16 + r; }
__device__ __forceinline__ int mfma16_row(double, int l16, int r) { return l16 + 4 * r; }

constexpr int kRhoMfmaTpw = 3;  // tiles per wave (a block: 2 x 3 tiles)

template <typename T, typename T2>
struct RhoMfmaArgs {
    const double4* pos;  // positions, charge in w
    int n;
    Box box;
    int ld, kmax;
    int wlog;     // log2 of the atoms per stage: 4..7 (a multiple of 16)
    int cpb;      // stages per block
    int nt;       // column tiles
    int ngroups;  // tile groups (blocks per chunk)
    int nchunks;
    const int* cols;  // nt * 16: (iky + kmax) << 16 | (ikz + kmax)
    const int* kmap;  // nt * 256: k of output (ikx, col), or -1
    int nk;
    double2* partial;
};

// LDS entries per staged atom: X (16) + Ey (2 kmax + 1) + Ez (2 kmax + 1),
// plus one so that the stride is odd and the stores of one row (consecutive
// atoms, 8 or 16 bytes each) fall in distinct banks.
__host__ __device__ inline int rho_mfma_stride(int kmax) {
    return 16 + 2 * (2 * kmax + 1) + 1;
}
// LDS bytes of one block: the staged rows and charges, or the sub-wave
// tiles, whichever is larger (the second reuses the first).
template <typename T>
__host__ __device__ inline size_t rho_mfma_lds(int width, int kmax) {
    size_t stage = size_t(width) * rho_mfma_stride(kmax) * 2 * sizeof(T) + size_t(width) * sizeof(T);
    size_t red = size_t(2) * 3 * kRhoMfmaTpw * 2 * 4 * 64 * sizeof(T);
    return stage > red ? stage : red;
}

template <typename T, typename T2>
__device__ __forceinline__ void rho_mfma_block(const RhoMfmaArgs<T, T2>& ra, int L, T2* __restrict__ smem) {
    typedef typename MfmaAcc<T>::type acc_t;
    int grp, chunk;
    if (!xcd_block_at(L, ra.ngroups, ra.nchunks, grp, chunk)) return;
    const int kmax = ra.kmax, wl = ra.wlog, W = 1 << wl, nt = ra.nt;
    constexpr int tpw = kRhoMfmaTpw;
    const int ny = 2 * kmax + 1, S = rho_mfma_stride(kmax);
    T* qs = reinterpret_cast<T*>(smem + W * S);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int sub = w & 3, half = w >> 2;
    const int kk = lane >> 4, c = lane & 15, comp = kk & 1;
    const int t0 = grp * 2 * tpw + half * tpw;
    int offy[kRhoMfmaTpw], offz[kRhoMfmaTpw];
#pragma unroll
    for (int t = 0; t < kRhoMfmaTpw; t++) {
        const int tt = t0 + t;
        // tiles past the last compute on a stand-in column and are not stored
        const int ci = tt < nt ? ra.cols[tt * 16 + c] : ((kmax << 16) | kmax);
        offy[t] = 16 + (ci >> 16);
        offz[t] = 16 + ny + (ci & 0xffff);
    }
    // the k-vector of each output, fetched now so that its latency is hidden
    int kmp[kRhoMfmaTpw][4];
#pragma unroll
    for (int t = 0; t < kRhoMfmaTpw; t++)
#pragma unroll
        for (int r = 0; r < 4; r++)
            kmp[t][r] = (sub == 0 && t0 + t < nt) ? ra.kmap[(t0 + t) * 256 + mfma16_row(T(0), kk, r) * 16 + c] : -1;
    acc_t are[kRhoMfmaTpw], aim[kRhoMfmaTpw];
#pragma unroll
    for (int t = 0; t < kRhoMfmaTpw; t++) {
        are[t] = acc_t{0, 0, 0, 0};
        aim[t] = acc_t{0, 0, 0, 0};
    }
    const int per = W >> 2;
    // staging: thread (axis, atom) forms the atom's phase factors along one
    // axis by the reference's recursion in FP64 (phase_axis64; the mixed form
    // rounds each factor to FP32 once), and the x-axis thread its charge;
    // atoms past n are zero
    for (int stg = 0; stg < ra.cpb; stg++) {
        const int a0 = (chunk * ra.cpb + stg) * W;
        if (stg > 0) __syncthreads();  // the previous stage is read before it is replaced
        if ((int)threadIdx.x < 3 * W) {
            const int a = threadIdx.x & (W - 1), axis = threadIdx.x >> wl, ga = a0 + a;
            double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
            if (ga < ra.n) p = ra.pos[ga];
            const double x = axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
            const double phase = (2.0 * kPi * ra.box.inv[axis]) * x;
            T2* row = smem + a * S + (axis == 0 ? 0 : (axis == 1 ? 16 + kmax : 16 + ny + kmax));
            T2 one;
            one.x = 1;
            one.y = 0;
            row[0] = one;
            double sn, cs;
            sincos(phase, &sn, &cs);
            double ex = cs, ey = sn;
            // x rows ikx = 0..kmax-1 (the half-space set), zero up to 15
            const int top = axis == 0 ? kmax - 1 : kmax;
            for (int k = 1; k <= top; k++) {
                if (k > 1) {
                    const double rx = ex * cs - ey * sn, ry = ex * sn + ey * cs;
                    ex = rx;
                    ey = ry;
                }
                T2 e, cj;
                e.x = cj.x = (T)ex;
                e.y = (T)ey;
                cj.y = -e.y;
                row[k] = e;
                if (axis != 0) row[-k] = cj;
            }
            if (axis == 0) {
                T2 zero;
                zero.x = zero.y = 0;
                for (int k = kmax > 1 ? kmax : 1; k < 16; k++) row[k] = zero;
                qs[a] = (T)p.w;
            }
        }
        __syncthreads();
        const int abase = sub * per + (kk >> 1);
        const T2* at0 = smem + abase * S;
        // per = width / 4 is a multiple of 4: two atom pairs per trip
        for (int p = 0; p < per; p += 4) {
#pragma unroll
            for (int u = 0; u < 4; u += 2) {
                const T2* at = at0 + (p + u) * S;
                const T2 x = at[c];
                const T qa = qs[abase + p + u];
                const T ar = qa * (comp ? -x.y : x.x);
                const T ai = qa * (comp ? x.x : x.y);
#pragma unroll
                for (int t = 0; t < kRhoMfmaTpw; t++) {
                    const T2 y = at[offy[t]], z = at[offz[t]];
                    const T b = comp ? (y.x * z.y + y.y * z.x) : (y.x * z.x - y.y * z.y);
                    are[t] = mfma16x16x4(ar, b, are[t]);
                    aim[t] = mfma16x16x4(ai, b, aim[t]);
                }
            }
        }
    }
    __syncthreads();  // staging reads done: the LDS now holds sub-wave tiles
    T* red = reinterpret_cast<T*>(smem);
    auto slot = [&](int s, int t, int part, int r) { return ((((half * 3 + s) * kRhoMfmaTpw + t) * 2 + part) * 4 + r) * 64 + lane; };
    if (sub > 0) {
#pragma unroll
        for (int t = 0; t < kRhoMfmaTpw; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                red[slot(sub - 1, t, 0, r)] = are[t][r];
                red[slot(sub - 1, t, 1, r)] = aim[t][r];
            }
    }
    __syncthreads();
    if (sub == 0) {
#pragma unroll
        for (int t = 0; t < kRhoMfmaTpw; t++) {
            const int tt = t0 + t;
            if (tt >= nt) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int k = kmp[t][r];
                double sr = (double)are[t][r], si = (double)aim[t][r];
                for (int s2 = 0; s2 < 3; s2++) {
                    sr += (double)red[slot(s2, t, 0, r)];
                    si += (double)red[slot(s2, t, 1, r)];
                }
                if (k >= 0) ra.partial[(size_t)chunk * ra.nk + k] = make_double2(sr, si);
            }
        }
    }
}

template <typename T, typename T2>
__global__ __launch_bounds__(512) void k_rho_mfma(RhoMfmaArgs<T, T2> ra) {
    extern __shared__ __align__(16) unsigned char rho_smem[];
    rho_mfma_block<T, T2>(ra, blockIdx.x, reinterpret_cast<T2*>(rho_smem));
}

// k_cells_rho with the matrix-core rho blocks; both parts use the one
// dynamic LDS buffer (max of sizeof(CellsLds) and rho_mfma_lds).
template <bool kMinImage, typename T, typename T2>
__global__ __launch_bounds__(kCellBlock) void k_cells_rho_mfma(CellsArgs ca, RhoMfmaArgs<T, T2> ra, int ncb) {
    extern __shared__ __align__(16) unsigned char rho_smem[];
    if ((int)blockIdx.x < ncb)
        nlist_cells_block<kMinImage>(ca, blockIdx.x, ncb, *reinterpret_cast<CellsLds*>(rho_smem));
    else
        rho_mfma_block<T, T2>(ra, blockIdx.x - ncb, reinterpret_cast<T2*>(rho_smem));
}

// k-space forces: one lane per atom, rows of k-vectors (fixed ikx, iky;
// consecutive ikz) read uniformly; F_i += q_i / (4 pi eps0) *
// sum_k Im(e^{ik.r_i} conj(rho_k)) field_k  (ewald.rs:699-719).
struct KRow {
    short ikx, iky, ikz0, count;  // (ikx, iky), first ikz, k-vectors in the row
    int kbase, pad;               // index of the row's first k-vector
};

// Block = 64 atoms x 4 waves.  grid.y splits the rows into ksplit ranges and
// wave w of the block takes the w-th quarter of its range; the four wave
// partials are added in order in LDS and split g writes fk[g * n + i]
// (combined in a fixed order by the consumer).  A lane reads only its own
// atom's column of the phase table, so the reads of a wave are contiguous
// and come from L2; rows have consecutive ikz.
__device__ __forceinline__ void kspace_wave_rows(int nrows, int split, int nsplit, int& r0, int& r1) {
    // wave index made visibly uniform, so that rows[] and the k records are
    // read with scalar loads
    int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int s0 = (int)(((long)nrows * split) / nsplit);
    const int s1 = (int)(((long)nrows * (split + 1)) / nsplit);
    r0 = s0 + ((s1 - s0) * w) / 4;
    r1 = s0 + ((s1 - s0) * (w + 1)) / 4;
}

__device__ __forceinline__ void kspace_block_store(double fx, double fy, double fz, int n, int ablock, int split,
                                                   double4* __restrict__ fk) {
    __shared__ double red[3][4][64];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    red[0][w][lane] = fx;
    red[1][w][lane] = fy;
    red[2][w][lane] = fz;
    __syncthreads();
    int i = ablock * 64 + lane;
    if (w == 0 && i < n) {
        double sx = ((red[0][0][lane] + red[0][1][lane]) + red[0][2][lane]) + red[0][3][lane];
        double sy = ((red[1][0][lane] + red[1][1][lane]) + red[1][2][lane]) + red[1][3][lane];
        double sz = ((red[2][0][lane] + red[2][1][lane]) + red[2][2][lane]) + red[2][3][lane];
        fk[(size_t)split * n + i] = make_double4(sx, sy, sz, 0.0);
    }
}

// One step along a row: s = Im(e conj(rho)), A += s g, B += s gz, e *= Ez[1].
template <typename T, typename T2, typename R>
__device__ __forceinline__ void kforce_step(const R& r, T2& e, T2 ez1, T& A, T& B) {
    T sv = e.y * r.re - e.x * r.im;
    A += sv * r.g;
    B += sv * r.gz;
    e = cmul(e, ez1);
}

// T = double: FP64 form; T = float: mixed form (FP32 phases, rho and
// factors; a wave's rows, a tenth of the k-vectors or fewer, are summed in
// FP32 and its sum joins the FP64 sum of the block's waves).  Records are
// read four at a time with one batch of scalar loads, so that a wave waits
// once per four k-vectors (a copy of the block's records in LDS measured no
// faster).  APL atoms per lane (block = 64 APL atoms): the row headers,
// records and their waits are shared by APL independent recursions.
template <typename T, typename T2, typename R, int APL>
__global__ __launch_bounds__(256) void k_kspace_force(const T2* __restrict__ tab, int ld, int n, int kmax,
                                                      const KRow* __restrict__ rows, int nrows,
                                                      const R* __restrict__ rec, double gx, double gy, double gz,
                                                      int nsplit, int nab, double4* __restrict__ fk) {
    int split, ablock;  // all splits of an atom block on one XCD
    if (!xcd_block(nsplit, nab, split, ablock)) return;
    const int lane = threadIdx.x & 63;
    int r0, r1;
    kspace_wave_rows(nrows, split, nsplit, r0, r1);
    const T2* col[APL];
    T2 ez1[APL];
#pragma unroll
    for (int u = 0; u < APL; u++) {
        int i = (ablock * APL + u) * 64 + lane;
        col[u] = tab + (i < ld ? i : ld - 1);
        // e^{i k.r} along a row: (Ex * Ey) * Ez[ikz0], then one product by
        // e^{i kz1 z} per step of ikz
        ez1[u] = col[u][(size_t)phase_row(2, 1, kmax) * ld];
    }
    const T tgx = (T)gx, tgy = (T)gy, tgz = (T)gz;
    T fx[APL], fy[APL], fz[APL];
    // software pipeline: the next row's header and phase columns are loaded
    // while the current row is summed
    KRow nrow = rows[r0 < r1 ? r0 : 0];
    T2 pex[APL], pey[APL], pez[APL];
#pragma unroll
    for (int u = 0; u < APL; u++) {
        fx[u] = fy[u] = fz[u] = 0.0;
        pex[u] = col[u][(size_t)phase_row(0, nrow.ikx, kmax) * ld];
        pey[u] = col[u][(size_t)phase_row(1, nrow.iky, kmax) * ld];
        pez[u] = col[u][(size_t)phase_row(2, nrow.ikz0, kmax) * ld];
    }
    // A full vector-memory wait before the row loop.  Without it the
    // compiler issues the Ez[1] load after the first row's columns and then
    // places a full wait inside the k loop (for Ez[1] in the first row),
    // which also drains the next row's prefetch at the start of every row;
    // with it, the loop waits for the columns once, at the top of the first
    // row, and each row's prefetch stays in flight while the row is summed.
    // 0x0F70: vmcnt(0), expcnt and lgkmcnt not waited on.
    __builtin_amdgcn_s_waitcnt(0x0F70);
    // The first row's columns go through an empty asm statement, so that
    // the values entering the row loop are not all loads: the compiler then
    // keeps the prefetch in place and in the same registers from row to row
    // (with a conditional prefetch it copied six registers in and out of
    // every row).  The prefetch is unconditional: the last row reloads its
    // own columns (in bounds, not used).
#pragma unroll
    for (int u = 0; u < APL; u++)
        asm volatile("" : "+v"(pex[u].x), "+v"(pex[u].y), "+v"(pey[u].x), "+v"(pey[u].y), "+v"(pez[u].x),
                     "+v"(pez[u].y));
    for (int r = r0; r < r1; r++) {
        KRow row = nrow;
        T2 e[APL];
        T A[APL], B[APL];
#pragma unroll
        for (int u = 0; u < APL; u++) {
            e[u] = cmul(cmul(pex[u], pey[u]), pez[u]);
            A[u] = 0;
            B[u] = 0;
        }
        nrow = rows[r + 1 < r1 ? r + 1 : r];
#pragma unroll
        for (int u = 0; u < APL; u++) {
            pex[u] = col[u][(size_t)phase_row(0, nrow.ikx, kmax) * ld];
            pey[u] = col[u][(size_t)phase_row(1, nrow.iky, kmax) * ld];
            pez[u] = col[u][(size_t)phase_row(2, nrow.ikz0, kmax) * ld];
        }
        const R* rp = rec + row.kbase;
        const int cnt = row.count;
        int c = 0;
        for (; c + 4 <= cnt; c += 4) {
            const R q0 = rp[c], q1 = rp[c + 1], q2 = rp[c + 2], q3 = rp[c + 3];
#pragma unroll
            for (int u = 0; u < APL; u++) {
                kforce_step(q0, e[u], ez1[u], A[u], B[u]);
                kforce_step(q1, e[u], ez1[u], A[u], B[u]);
                kforce_step(q2, e[u], ez1[u], A[u], B[u]);
                kforce_step(q3, e[u], ez1[u], A[u], B[u]);
            }
        }
        for (; c < cnt; c++) {
            const R q0 = rp[c];
#pragma unroll
            for (int u = 0; u < APL; u++) kforce_step(q0, e[u], ez1[u], A[u], B[u]);
        }
#pragma unroll
        for (int u = 0; u < APL; u++) {
            fx[u] += A[u] * (tgx * (T)row.ikx);
            fy[u] += A[u] * (tgy * (T)row.iky);
            fz[u] += B[u] * tgz;
        }
    }
#pragma unroll
    for (int u = 0; u < APL; u++) {
        if (u > 0) __syncthreads();  // the LDS of the previous store is read before it is reused
        kspace_block_store((double)fx[u], (double)fy[u], (double)fz[u], n, ablock * APL + u, split, fk);
    }
}

// ---- mixed-form k-space ----------------------------------------------------------
// Phase factors in FP32 from phases range-reduced in FP64: for every n the
// phase n k_axis x is formed and reduced to [-pi, pi] in FP64, then
// sincosf.  Same row layout as k_phase.
__global__ void k_phase32(const double4* __restrict__ pos, int n, int ld, Box b, int kmax, float2* __restrict__ tab,
                          float* __restrict__ q) {
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    int axis = gid / ld;
    int a = gid - axis * ld;
    if (axis >= 3) return;
    bool live = a < n;
    double x = 0.0, qa = 0.0;
    if (live) {
        double4 p = pos[a];
        x = axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
        qa = p.w;
    }
    phase_axis32(x, b.inv[axis], axis, kmax, tab, ld, a);
    if (axis == 0) q[a] = live ? (float)qa : 0.f;
}

// ---------------------------------------------------------------------------------
// Integrators (lumol-sim/src/md/integrators.rs), FP64, unfused as in the
// reference so that positions follow the reference's rounding.
// ---------------------------------------------------------------------------------
#pragma clang fp contract(off)
// Neighbour-list displacement test of the new position.
__device__ __forceinline__ bool moved_beyond(double4 p, const double4* __restrict__ ref, int i, double half_skin2) {
    double4 r = ref[i];
    double dx = p.x - r.x, dy = p.y - r.y, dz = p.z - r.z;
    return dx * dx + dy * dy + dz * dz > half_skin2;
}

// Non-finite state of an atom after an integration step: flags[2] is set
// (wave-level OR, one vector atomic per wave) so that the host can fail the
// run instead of returning NaN positions as a result.
__device__ __forceinline__ bool nonfinite3(double x, double y, double z) {
    return !isfinite(x) || !isfinite(y) || !isfinite(z);
}
__device__ __forceinline__ void flag_nonfinite(bool bad, int* __restrict__ flags) {
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(flags + 2, 1);
}

// Berendsen factor (thermostats.rs:113-121) from the kinetic energy
// partials of the previous step, dof = 3N (system.rs:249-255): the block sums
// the partials in a fixed order (thread-strided, waves, then wave order), so
// every block of a launch, and k_berendsen, get the same bits.
struct Thermo {
    const double* ke_part;  // null: no pending scale
    int nparts;
    double dof, T0, tau;
};

// The sum runs over the first 256 threads whatever the block size, so that
// every kernel computes the same lambda bit for bit.
__device__ __forceinline__ double berendsen_lambda(const Thermo& th) {
    __shared__ double red[4];
    __shared__ double lam;
    double t = 0.0;
    if (threadIdx.x < 256)
        for (int p = threadIdx.x; p < th.nparts; p += 256) t += th.ke_part[p];
    t = wave_sum(t);
    if ((threadIdx.x & 63) == 0 && threadIdx.x < 256) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double kin = red[0] + red[1] + red[2] + red[3];
        double T = 2.0 * kin / (th.dof * kBoltzmann);
        lam = sqrt(1.0 + (th.T0 / T - 1.0) / th.tau);
    }
    __syncthreads();
    return lam;
}

// Phase-table part of a first half-step: role r (1..3) of the block writes
// the rows of axis r - 1 for the block's 64 new positions, staged in LDS by
// role 0 (charges with axis 0).  Padding atoms (n <= i < ld) get x = 0 and
// q = 0, as k_phase.
// FP64 table: roles 1-3, one axis each (the recursion runs along k).
// FP32 table (every k reduced and evaluated on its own): with 8 waves per
// block, roles 1-6 write (axis, half of the k range), role 7 idles.
__device__ __forceinline__ void phase_roles(const PhaseOut& ph, int i, int n, int role, int lane,
                                            const double (&px)[4][64]) {
    if (role == 0 || i >= ph.ld) return;
    if (ph.mode == 1) {
        if (role > 3) return;
        const int axis = role - 1;
        const double x = i < n ? px[axis][lane] : 0.0;
        phase_axis64(x, ph.box.inv[axis], axis, ph.kmax, ph.tab, ph.ld, i);
        if (axis == 0) ph.q[i] = i < n ? px[3][lane] : 0.0;
    } else {
        const int nsp = ((int)blockDim.x / 64 - 1) / 3;  // 1 (4 waves) or 2 (8 waves)
        const int r = role - 1;
        if (r >= 3 * nsp) return;
        const int axis = r % 3, part = r / 3;
        const int k0 = part * ph.kmax / nsp + 1, k1 = (part + 1) * ph.kmax / nsp;
        const double x = i < n ? px[axis][lane] : 0.0;
        phase_axis32(x, ph.box.inv[axis], axis, ph.kmax, ph.tab32, ph.ld, i, k0, k1);
        if (axis == 0 && part == 0) ph.q32[i] = i < n ? (float)px[3][lane] : 0.f;
    }
}

// VelocityVerlet first half (integrators.rs:47-53): the pending Berendsen
// scale of the previous step (th.ke_part != null), then v += (dt/2) a;
// x += v dt; the displacement test of the neighbour list (sets flags[0]);
// and the phase table of the new position when `ph.mode` != 0.
// Block = 64 atoms x 4 roles (FP64 table) or 8 roles (FP32 table): role 0
// integrates, the others write the phase rows (phase_roles), so that the
// trigonometry of an atom runs on three or six waves.  Grid:
// ceil(max(n, ld) / 64).
__global__ __launch_bounds__(512) void k_vv_first(double4* __restrict__ pos, double4* __restrict__ vel,
                                                  const double4* __restrict__ acc, int n, double dt, Thermo th,
                                                  const double4* __restrict__ ref, double half_skin2,
                                                  int* __restrict__ flags, PhaseOut ph, CellTag tg) {
    __shared__ double px[4][64];
    const int lane = threadIdx.x & 63, role = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    double l = th.ke_part ? berendsen_lambda(th) : 1.0;
    if (role == 0) {
        bool moved = false, bad = false;
        double4 p = make_double4(0, 0, 0, 0);
        if (i < n) {
            double4 v = vel[i];
            if (th.ke_part) {
                v.x = v.x * l;
                v.y = v.y * l;
                v.z = v.z * l;
            }
            double4 a = acc[i];
            double h = 0.5 * dt;
            v.x += h * a.x;
            v.y += h * a.y;
            v.z += h * a.z;
            p = pos[i];
            p.x += v.x * dt;
            p.y += v.y * dt;
            p.z += v.z * dt;
            vel[i] = v;
            pos[i] = p;
            moved = moved_beyond(p, ref, i, half_skin2);
            bad = nonfinite3(p.x, p.y, p.z) || nonfinite3(v.x, v.y, v.z);
            if (tg.tag) tg.tag[i] = cell_tag(tg, p);
            if (tg.frac) tg.frac[i] = frac_pos(tg.b, tg.fo, p);
        }
        if (__any(moved) && lane == 0) atomicOr(flags, 1);
        flag_nonfinite(bad, flags);
        px[0][lane] = p.x;
        px[1][lane] = p.y;
        px[2][lane] = p.z;
        px[3][lane] = p.w;
    }
    if (!ph.mode) return;  // uniform: a kernel argument
    __syncthreads();
    phase_roles(ph, i, n, role, lane, px);
}

// LeapFrog (integrators.rs:150-168), position part, after the pending
// Berendsen scale of the previous step, with the displacement test and the
// phase table as k_vv_first (same block layout).
__global__ __launch_bounds__(512) void k_lf_first(double4* __restrict__ pos, double4* __restrict__ vel,
                                                  const double4* __restrict__ acc, int n, double dt, Thermo th,
                                                  const double4* __restrict__ ref, double half_skin2,
                                                  int* __restrict__ flags, PhaseOut ph, CellTag tg) {
    __shared__ double px[4][64];
    const int lane = threadIdx.x & 63, role = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    double l = th.ke_part ? berendsen_lambda(th) : 1.0;
    if (role == 0) {
        bool moved = false, bad = false;
        double4 p = make_double4(0, 0, 0, 0);
        if (i < n) {
            double4 v = vel[i];
            if (th.ke_part) {
                v.x = v.x * l;
                v.y = v.y * l;
                v.z = v.z * l;
                vel[i] = v;
            }
            double4 a = acc[i];
            double dt2 = dt * dt;
            p = pos[i];
            p.x += v.x * dt + (0.5 * a.x) * dt2;
            p.y += v.y * dt + (0.5 * a.y) * dt2;
            p.z += v.z * dt + (0.5 * a.z) * dt2;
            pos[i] = p;
            moved = moved_beyond(p, ref, i, half_skin2);
            bad = nonfinite3(p.x, p.y, p.z) || nonfinite3(v.x, v.y, v.z);
            if (tg.tag) tg.tag[i] = cell_tag(tg, p);
            if (tg.frac) tg.frac[i] = frac_pos(tg.b, tg.fo, p);
        }
        if (__any(moved) && lane == 0) atomicOr(flags, 1);
        flag_nonfinite(bad, flags);
        px[0][lane] = p.x;
        px[1][lane] = p.y;
        px[2][lane] = p.z;
        px[3][lane] = p.w;
    }
    if (!ph.mode) return;
    __syncthreads();
    phase_roles(ph, i, n, role, lane, px);
}

// Second half of VelocityVerlet (integrators.rs:55-68) or LeapFrog
// (:160-167).  The total force is the pair/intra-molecular force plus the
// k-space partials q_i / (4 pi eps0) sum_g fk[g][i] (ewald.rs:716-719).
// With a thermostat (ke_part != null), block partials of the kinetic energy
// are written for k_berendsen.  Blocks of 64 atoms (one wave), so that the
// loads of a small system spread over many CUs.
template <bool kLeapFrog>
__global__ __launch_bounds__(64) void k_md_second(const double4* __restrict__ force, const double4* __restrict__ pos,
                                                   const double4* __restrict__ contrib,
                                                   const int* __restrict__ slot_start,
                                                   const double4* __restrict__ fk, int ksplit,
                                                   double4* __restrict__ vel, double4* __restrict__ acc, int n,
                                                   double dt, double* __restrict__ ke_part,
                                                   int* __restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0;
    bool bad = false;
    if (i < n) {
        double4 f = total_force(i, n, force, contrib, slot_start, fk, ksplit, pos);
        double4 v = vel[i];
        double m = v.w;
        double4 na = make_double4(f.x / m, f.y / m, f.z / m, 0.0);
        if (kLeapFrog) {
            double4 a = acc[i];
            v.x += (0.5 * (a.x + na.x)) * dt;
            v.y += (0.5 * (a.y + na.y)) * dt;
            v.z += (0.5 * (a.z + na.z)) * dt;
        } else {
            double h = 0.5 * dt;
            v.x += h * na.x;
            v.y += h * na.y;
            v.z += h * na.z;
        }
        acc[i] = na;
        vel[i] = v;
        ke = 0.5 * m * (v.x * v.x + v.y * v.y + v.z * v.z);
        bad = nonfinite3(v.x, v.y, v.z) || nonfinite3(na.x, na.y, na.z);
    }
    flag_nonfinite(bad, flags);
    if (!ke_part) return;
    double s = wave_sum(ke);
    if (threadIdx.x == 0) ke_part[blockIdx.x] = s;
}

// The second half of step k and the first half of step k + 1 in one launch
// (inside lg_md_run, without a thermostat: the first half then needs no
// global kinetic energy).  The arithmetic of k_md_second followed by
// k_vv_first / k_lf_first, in the same order, on the values those kernels
// pass through memory: the same bits, one launch fewer per step.  Block
// layout of the first-half kernels (64 atoms; the other waves write the
// phase table).
template <bool kLeapFrog>
__global__ __launch_bounds__(512) void k_md_fused(const double4* __restrict__ force, double4* __restrict__ pos,
                                                  const double4* __restrict__ contrib,
                                                  const int* __restrict__ slot_start,
                                                  const double4* __restrict__ fk, int ksplit,
                                                  double4* __restrict__ vel, double4* __restrict__ acc, int n,
                                                  double dt, const double4* __restrict__ ref, double half_skin2,
                                                  int* __restrict__ flags, PhaseOut ph, CellTag tg) {
    __shared__ double px[4][64];
    const int lane = threadIdx.x & 63, role = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    if (role == 0) {
        bool moved = false, bad = false;
        double4 p = make_double4(0, 0, 0, 0);
        if (i < n) {
            // second half of step k
            double4 f = total_force(i, n, force, contrib, slot_start, fk, ksplit, pos);
            double4 v = vel[i];
            double m = v.w;
            double4 na = make_double4(f.x / m, f.y / m, f.z / m, 0.0);
            if (kLeapFrog) {
                double4 a = acc[i];
                v.x += (0.5 * (a.x + na.x)) * dt;
                v.y += (0.5 * (a.y + na.y)) * dt;
                v.z += (0.5 * (a.z + na.z)) * dt;
            } else {
                double h = 0.5 * dt;
                v.x += h * na.x;
                v.y += h * na.y;
                v.z += h * na.z;
            }
            acc[i] = na;
            // first half of step k + 1
            p = pos[i];
            if (kLeapFrog) {
                double dt2 = dt * dt;
                p.x += v.x * dt + (0.5 * na.x) * dt2;
                p.y += v.y * dt + (0.5 * na.y) * dt2;
                p.z += v.z * dt + (0.5 * na.z) * dt2;
            } else {
                double h = 0.5 * dt;
                v.x += h * na.x;
                v.y += h * na.y;
                v.z += h * na.z;
                p.x += v.x * dt;
                p.y += v.y * dt;
                p.z += v.z * dt;
            }
            vel[i] = v;
            pos[i] = p;
            moved = moved_beyond(p, ref, i, half_skin2);
            bad = nonfinite3(p.x, p.y, p.z) || nonfinite3(v.x, v.y, v.z);
            if (tg.tag) tg.tag[i] = cell_tag(tg, p);
            if (tg.frac) tg.frac[i] = frac_pos(tg.b, tg.fo, p);
        }
        if (__any(moved) && lane == 0) atomicOr(flags, 1);
        flag_nonfinite(bad, flags);
        px[0][lane] = p.x;
        px[1][lane] = p.y;
        px[2][lane] = p.z;
        px[3][lane] = p.w;
    }
    if (!ph.mode) return;  // uniform: a kernel argument
    __syncthreads();
    phase_roles(ph, i, n, role, lane, px);
}

// Berendsen factor of the last step of a run into *lambda (one block, same
// sum as berendsen_lambda).
__global__ __launch_bounds__(256) void k_berendsen(Thermo th, double* __restrict__ lambda) {
    double l = berendsen_lambda(th);
    if (threadIdx.x == 0) *lambda = l;
}

__global__ void k_scale_vel(double4* __restrict__ vel, int n, const double* __restrict__ lambda) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double l = *lambda;
    double4 v = vel[i];
    v.x = v.x * l;
    v.y = v.y * l;
    v.z = v.z * l;
    vel[i] = v;
}
#pragma clang fp contract(on)

__global__ void k_set_scalar(double* __restrict__ p, double v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

}  // namespace lg
