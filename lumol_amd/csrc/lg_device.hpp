// lg_device.hpp -- device-side parameter blocks and scalar physics for the
// gfx950 force-and-energy kernels.
//
// The scalar functions restate Lumol's formulas (reference checkout paths):
//   Lennard-Jones      lumol-core/src/energy/functions.rs:79-89
//   Harmonic           lumol-core/src/energy/functions.rs:143-151
//   other potentials   lumol-core/src/energy/functions.rs:198-207 (CosineHarmonic),
//                      :244-256 (Torsion), :286-300 (Buckingham),
//                      :353-367 (BornMayerHuggins), :414-424 (Morse),
//                      :474-482 (Gaussian), :552-566 (Mie)
//   TableComputation   lumol-core/src/energy/computations.rs:122-146
//   PairInteraction    lumol-core/src/energy/pairs.rs:185-218 (r >= rc -> 0)
//   minimum image      lumol-core/src/sys/config/cells.rs:284-300
//   Ewald real space   lumol-core/src/energy/global/ewald.rs:387-428 (r > rc -> 0)
//   Wolf               lumol-core/src/energy/global/wolf.rs:90-117 (r > rc -> 0)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lg {

constexpr double kPi = 3.14159265358979323846;
constexpr double kFrac2SqrtPi = 1.12837916709551257390;  // 2/sqrt(pi)
// lumol-core/src/consts.rs:9,15
constexpr double kBoltzmann = 8.31446284161522e-7;
constexpr double kFourPiEps0 = 7.197589831304046;

enum : int {
    POT_NONE = -1,
    POT_NULL = 0,
    POT_LJ = 1,
    POT_HARMONIC = 2,
    POT_BUCKINGHAM = 3,
    POT_BORN = 4,
    POT_MORSE = 5,
    POT_GAUSSIAN = 6,
    POT_MIE = 7,
    POT_COSINE_HARMONIC = 8,
    POT_TORSION = 9
};
// restrictions.rs:37-50
enum : int { BP_NONE = 0, BP_SAME = 1, BP_ONE = 2, BP_TWO = 3, BP_THREE = 4, BP_FAR = 5 };
enum : int { COUL_NONE = 0, COUL_EWALD = 1, COUL_WOLF = 2 };

// One entry per (kind_i, kind_j), symmetric, K*K entries.  Parameters in
// the reference's field order: LJ {sigma, epsilon}; Harmonic {k, x0};
// Buckingham {a, c, rho}; BornMayerHuggins {a, c, d, sigma, rho};
// Morse {a, x0, depth}; Gaussian {a, b}; Mie {sigma, n, m, prefac}.
// tab_n > 0: a TableComputation of the potential, whose energies are
// tabs[tab_off + i] and forces tabs[tab_off + tab_n + i] at r = i tab_delta.
struct PairParam {
    int pot;           // POT_* ; POT_NONE when no interaction is set
    int inter_excl;    // 1 when the restriction excludes inter-molecular pairs
    double p0, p1;
    double rc;         // PairInteraction cutoff
    double shift;      // V(rc) when shifted, else 0
    double p2, p3, p4;
    int tab_n, tab_off;
    double tab_delta;
};

// The fields of PairParam that the LJ/Harmonic/Null pair kernels read
// (the table they stage in LDS).
struct PairFast {
    int pot, inter_excl;
    double p0, p1, rc, shift;
};
// Kinds the fast pair kernels evaluate; a system with another kind or a
// table uses the general pair kernel.
__host__ __device__ __forceinline__ bool pot_is_fast(int pot, int tab_n) {
    return tab_n == 0 && pot <= POT_HARMONIC;
}
// Marks, in a PairFast, an entry that needs the general evaluation from the
// full PairParam (the Monte Carlo kernels, which run every kind).
constexpr int POT_GENERAL = 64;
__host__ __device__ __forceinline__ PairFast pair_fast(const PairParam& p) {
    return PairFast{pot_is_fast(p.pot, p.tab_n) ? p.pot : POT_GENERAL, p.inter_excl, p.p0, p.p1, p.rc, p.shift};
}

// Bond, angle and dihedral potential: Harmonic {k, x0}, Morse {a, x0, depth},
// CosineHarmonic {k, cos(x0)}, Torsion {k, delta, n}, Null.
struct BondedParam {
    int pot;  // POT_* (POT_NONE: no potential)
    int tab_n, tab_off;  // always 0 (no tables for bonded terms)
    double p0, p1, p2, p3, p4;
    double tab_delta;
};

struct CoulParam {
    int kind;          // COUL_*
    int inter_excl;    // restriction excludes inter-molecular pairs
    double alpha, rc;
    double e_const, f_const;  // Wolf energy_constant / force_constant
};

// Orthorhombic cell: lengths and reciprocal lengths.
struct Box {
    double L[3];
    double inv[3];
    int periodic;
};

// ---- geometry -----------------------------------------------------------------
// Minimum image of d = r_i - r_j, bit-identical to
// `d -= round(d / a) * a` (cells.rs:288-290).  The quotient is formed with the
// reciprocal and re-done with an IEEE division only when it lies within a few
// ulps of a half-integer, the only case where the two roundings can differ.
#pragma clang fp contract(off)
__device__ __forceinline__ double image1(double d, double L, double inv) {
    double q = d * inv;
    double n = round(q);
    double frac = fabs(q - trunc(q));
    if (fabs(frac - 0.5) < 1e-9) n = round(d / L);
    return d - n * L;
}

__device__ __forceinline__ void min_image(const Box& b, double& dx, double& dy, double& dz) {
    if (b.periodic) {
        dx = image1(dx, b.L[0], b.inv[0]);
        dy = image1(dy, b.L[1], b.inv[1]);
        dz = image1(dz, b.L[2], b.inv[2]);
    }
}

// Minimum image for the mixed form, whose distance is rounded to FP32 next:
// d - rint(d / a) a with the reciprocal, 3 operations per axis.  It differs
// from image1 only for |d| within ulps of a half box length, where both
// images are equally near and, the cutoff being at most half the box, no
// pair within it lies.
__device__ __forceinline__ void min_image_fast(const Box& b, double& dx, double& dy, double& dz) {
    if (b.periodic) {
        dx -= rint(dx * b.inv[0]) * b.L[0];
        dy -= rint(dy * b.inv[1]) * b.L[1];
        dz -= rint(dz * b.inv[2]) * b.L[2];
    }
}

// Vector3D::norm2 order: x*x + y*y + z*z, left to right, unfused.
__device__ __forceinline__ double norm2_ordered(double x, double y, double z) {
    double a = x * x;
    double b = y * y;
    double c = z * z;
    return (a + b) + c;
}
#pragma clang fp contract(on)

// ---- pair potentials (FP64) ----------------------------------------------------------
// energy and force / r  (force() returns -dV/dr; the vector force is (f/r) d).
__device__ __forceinline__ void pot_eval(int pot, double p0, double p1, double r, double inv_r,
                                         double& e, double& f_over_r) {
    if (pot == POT_LJ) {
        double x = p0 * inv_r;
        double x2 = x * x;
        double s6 = x2 * x2 * x2;
        e = 4.0 * p1 * (s6 * s6 - s6);
        f_over_r = -24.0 * p1 * (s6 - 2.0 * (s6 * s6)) * inv_r * inv_r;
    } else if (pot == POT_HARMONIC) {
        double dx = r - p1;
        e = 0.5 * p0 * dx * dx;
        f_over_r = p0 * (p1 - r) * inv_r;
    } else {
        e = 0.0;
        f_over_r = 0.0;
    }
}

// Every potential kind and TableComputation: energy and force / r (with
// inv_r = 1 the plain force, -dV/dx, for angles and dihedrals).  Operation
// order as the reference; P is PairParam or BondedParam.
template <class P>
__device__ __forceinline__ void pot_eval_any(const P& q, const double* __restrict__ tabs, double r, double inv_r,
                                             double& e, double& f_over_r) {
    if (q.tab_n > 0) {
        // computations.rs:122-146: `floor(r / delta) as usize` saturates at 0
        const double u = floor(r / q.tab_delta);
        const double bin = u > 0.0 ? u : 0.0;  // NaN -> 0
        e = 0.0;
        double f = 0.0;
        if (bin < (double)(q.tab_n - 1)) {
            const int b = (int)bin;
            const double dx = r - bin * q.tab_delta;
            const double* te = tabs + q.tab_off;
            const double* tf = te + q.tab_n;
            e = te[b] + dx * ((te[b + 1] - te[b]) / q.tab_delta);
            f = tf[b] + dx * ((tf[b + 1] - tf[b]) / q.tab_delta);
        }
        f_over_r = f * inv_r;
        return;
    }
    double f = 0.0;
    switch (q.pot) {
        case POT_LJ: pot_eval(POT_LJ, q.p0, q.p1, r, inv_r, e, f_over_r); return;
        case POT_HARMONIC: pot_eval(POT_HARMONIC, q.p0, q.p1, r, inv_r, e, f_over_r); return;
        case POT_BUCKINGHAM: {
            const double r3 = r * r * r;
            const double ex = exp(-r / q.p2);
            e = q.p0 * ex - q.p1 / (r3 * r3);
            f = q.p0 / q.p2 * ex - 6.0 * q.p1 / (r3 * r3 * r);
            break;
        }
        case POT_BORN: {
            const double r2 = r * r;
            const double r6 = r2 * r2 * r2;
            const double r7 = r6 * r;
            const double ex = exp((q.p3 - r) / q.p4);
            e = q.p0 * ex - q.p1 / r6 + q.p2 / (r6 * r2);
            f = q.p0 / q.p4 * ex - 6.0 * q.p1 / r7 + 8.0 * q.p2 / (r7 * r2);
            break;
        }
        case POT_MORSE: {
            const double ex = exp((q.p1 - r) * q.p0);
            const double m = 1.0 - ex;
            e = q.p2 * m * m;
            f = 2.0 * q.p2 * (1.0 - ex * ex) * q.p0;
            break;
        }
        case POT_GAUSSIAN: {
            e = -q.p0 * exp(-q.p1 * r * r);
            f = 2.0 * q.p1 * r * e;
            break;
        }
        case POT_MIE: {
            const double sr = q.p0 / r;
            const double rep = pow(sr, q.p1), att = pow(sr, q.p2);
            e = q.p3 * (rep - att);
            f = q.p3 * (q.p1 * rep - q.p2 * att) / r;
            break;
        }
        case POT_COSINE_HARMONIC: {
            const double c = cos(r);
            const double dr = c - q.p1;
            e = 0.5 * q.p0 * dr * dr;
            f = q.p0 * dr * sin(r);
            break;
        }
        case POT_TORSION: {
            const double x = q.p2 * r - q.p1;
            e = q.p0 * (1.0 + cos(x));
            f = q.p0 * q.p2 * sin(x);
            break;
        }
        default: e = 0.0;
    }
    f_over_r = f * inv_r;
}

// Ewald real-space pair: energy and the factor multiplying d in the force
// vector (ewald.rs:387-428, force = force_pair(r) * rij at :490).  `excluded`
// selects the k-space removal term (-erf) used for excluded pairs.
__device__ __forceinline__ void ewald_real(double alpha, double qiqj, double r, double inv_r, bool excluded,
                                           double& e, double& f_over_r) {
    double ar = alpha * r;
    double ex = exp(-alpha * alpha * r * r);
    double g = alpha * kFrac2SqrtPi * ex;
    double pref = qiqj / kFourPiEps0;
    if (excluded) {
        double erfv = erf(ar);
        e = -pref * erfv * inv_r;
        f_over_r = pref * inv_r * inv_r * (g - erfv * inv_r);
    } else {
        double erfcv = erfc(ar);
        e = pref * erfcv * inv_r;
        f_over_r = pref * inv_r * inv_r * (g + erfcv * inv_r);
    }
}

// Wolf pair, energy and force/r (wolf.rs:90-117).
__device__ __forceinline__ void wolf_pair(const CoulParam& c, double qiqj, double r, double inv_r, double& e,
                                          double& f_over_r) {
    double ar = c.alpha * r;
    double erfcv = erfc(ar);
    double ex = exp(-ar * ar);
    e = qiqj * (erfcv * inv_r - c.e_const) / kFourPiEps0;
    double factor = erfcv * inv_r * inv_r + c.alpha * kFrac2SqrtPi * ex * inv_r;
    f_over_r = qiqj * (factor - c.f_const) * inv_r / kFourPiEps0;
}

// ---- restrictions (restrictions.rs:85-116) ------------------------------------------
__device__ __host__ __forceinline__ void restriction_info(int restriction, double scale14, int path,
                                                          bool& excluded, double& scaling) {
    bool same = path != BP_NONE;
    switch (restriction) {
        case 0: excluded = false; break;                            // None
        case 1: excluded = !same; break;                            // IntraMolecular
        case 2: excluded = same; break;                             // InterMolecular
        case 3: excluded = path == BP_ONE; break;                   // Exclude12
        case 4: case 6: excluded = path == BP_ONE || path == BP_TWO; break;  // Exclude13, Scale14
        case 5: excluded = path == BP_ONE || path == BP_TWO || path == BP_THREE; break;
        default: excluded = false;
    }
    scaling = (restriction == 6 && path == BP_THREE) ? scale14 : 1.0;
}

// ---- wave reductions -----------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace lg

// =====================================================================================
// Mixed-precision helpers (FP32 arithmetic on FP64-derived inputs)
// =====================================================================================
namespace lg {

// erfc(x) = exp(-x^2) erfcx(x) with erfcx from a piecewise polynomial table
// over [0, alpha rc]: kErfcxPieces equal pieces, each a polynomial in
// t in [-1, 1] (the piece's Chebyshev interpolant at degree kErfcxDeg64 or
// kErfcxDeg32, fitted on the host in long double against erfcl and checked
// there).  exp(-x^2) is the factor the Ewald/Wolf force needs anyway, so
// erfc costs a few FMAs instead of a second exponential and a rational
// approximation.
constexpr int kErfcxPieces = 256;
constexpr int kErfcxDeg64 = 5;
constexpr int kErfcxDeg32 = 3;
// The piece index is clamped in floating point before the conversion (NaN
// lands in piece 0), so that a non-finite distance reads inside the table.
__host__ __device__ __forceinline__ double erfcx_tab(const double* __restrict__ c, double x, double inv_h) {
    const double u = x * inv_h;
    const int k = (int)fmin(fmax(u, 0.0), (double)(kErfcxPieces - 1));
    const double t = 2.0 * (u - (double)k) - 1.0;
    const double* q = c + k * (kErfcxDeg64 + 1);
    double p = q[kErfcxDeg64];
#pragma unroll
    for (int d = kErfcxDeg64 - 1; d >= 0; d--) p = p * t + q[d];
    return p;
}
__host__ __device__ __forceinline__ float erfcx_tab(const float* __restrict__ c, float x, float inv_h) {
    const float u = x * inv_h;
    // x >= 0 (a distance or alpha times one), so only the top needs a
    // clamp; a NaN u gives the last piece (fminf returns the other operand)
    const int k = (int)fminf(u, (float)(kErfcxPieces - 1));
    const float t = 2.0f * (u - (float)k) - 1.0f;
    const float* q = c + k * (kErfcxDeg32 + 1);
    float p = q[kErfcxDeg32];
#pragma unroll
    for (int d = kErfcxDeg32 - 1; d >= 0; d--) p = p * t + q[d];
    return p;
}

// erfc(x) for x >= 0 from the Chebyshev fit of Numerical Recipes in C, 2nd
// ed., section 6.2 ("erfcc"): fractional error below 1.2e-7 everywhere.
// `ex2` = exp(-x^2) is returned for the Ewald force term.
__device__ __forceinline__ float erfc_cheb(float x, float& ex2) {
    float t = 1.0f / (1.0f + 0.5f * x);
    float p = -1.26551223f +
              t * (1.00002368f +
                   t * (0.37409196f +
                        t * (0.09678418f +
                             t * (-0.18628806f +
                                  t * (0.27886807f +
                                       t * (-1.13520398f +
                                            t * (1.48851587f + t * (-0.82215223f + t * 0.17087277f))))))));
    float mx2 = -x * x;
    ex2 = __expf(mx2);
    return t * __expf(mx2 + p);
}

}  // namespace lg
