// lumol_oracle.cpp -- CPU FP64 restatement of Lumol's force-and-energy path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the timed
// CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it; the product library never links or calls it.
//
// It restates, in plain C++17, the algorithm of the reference's Rust sources
// (paths relative to the reference checkout):
//   pair potentials      lumol-core/src/energy/functions.rs:30-107,128-160
//   PairInteraction      lumol-core/src/energy/pairs.rs:57-297
//   restrictions         lumol-core/src/energy/restrictions.rs:85-116
//   bond paths           lumol-core/src/sys/config/bonding.rs:78-155,
//                        lumol-core/src/sys/config/configuration.rs:124-144
//   minimum image        lumol-core/src/sys/config/cells.rs:263-320
//   angle derivatives    lumol-core/src/sys/config/cells.rs:335-359
//   Forces/energy/virial lumol-core/src/sys/compute.rs:30-255,
//                        lumol-core/src/sys/energy.rs:30-158
//   Ewald                lumol-core/src/energy/global/ewald.rs:115-183,278-378,
//                        387-530,619-733,756-839,888-914
//   Wolf                 lumol-core/src/energy/global/wolf.rs:68-283
//   Velocity-Verlet      lumol-sim/src/md/integrators.rs:22-70
//   LeapFrog             lumol-sim/src/md/integrators.rs:128-169
//   Berendsen thermostat lumol-sim/src/md/thermostats.rs:113-121
//   EnergyCache (MC)     lumol-core/src/sys/cache.rs:75-220
//
// Arithmetic is written in the reference's operation order and the file is
// compiled with -ffp-contract=off (Rust never fuses multiply-adds), so the
// scalar functions reproduce the reference's values bit for bit where libm's
// erf/erfc/exp agree with the `special` crate.  Parallel sums follow the
// reference's structure (per-i work, thread-local force buffers merged in a
// fixed order) so results are deterministic for a given thread count.
//
// Parity pinning: see the golden-vector tests in tests/test_oracle_golden.py.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <set>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

namespace {

constexpr double PI = 3.14159265358979323846;
constexpr double FRAC_2_SQRT_PI = 1.12837916709551257390;  // 2/sqrt(pi)
// lumol-core/src/consts.rs:9,15
constexpr double K_BOLTZMANN = 8.31446284161522e-7;
constexpr double FOUR_PI_EPSILON_0 = 7.197589831304046;

thread_local std::string g_error;

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    double operator[](int k) const { return k == 0 ? x : (k == 1 ? y : z); }
    double& operator[](int k) { return k == 0 ? x : (k == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3& operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline V3& operator-=(V3& a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
inline double norm2(V3 a) { return dot(a, a); }
inline double norm(V3 a) { return std::sqrt(norm2(a)); }

struct M3 {
    double m[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
};
inline M3& operator+=(M3& a, const M3& b) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) a.m[i][j] += b.m[i][j];
    return a;
}
// a.tensorial(b): (i,j) = a[i]*b[j]   (types/vectors.rs:150)
inline M3 tensorial(V3 a, V3 b) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = a[i] * b[j];
    return r;
}
inline M3 scale(double s, const M3& a) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = s * a.m[i][j];
    return r;
}

// --- complex numbers, types/complex.rs -------------------------------------
struct Cx {
    double re = 0, im = 0;
};
inline Cx cmul(Cx a, Cx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline Cx conj(Cx a) { return {a.re, -a.im}; }
inline Cx polar(double r, double phi) { return {r * std::cos(phi), r * std::sin(phi)}; }
inline double cnorm2(Cx a) { return a.re * a.re + a.im * a.im; }

// --- potentials ------------------------------------------------------------
enum PotKind {
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

// Parameters, in the reference's field order:
//   LJ {sigma, epsilon}; Harmonic {k, x0}; Buckingham {a, c, rho};
//   BornMayerHuggins {a, c, d, sigma, rho}; Morse {a, x0, depth};
//   Gaussian {a, b}; Mie {sigma, n, m, prefac} (prefac from Mie::new);
//   CosineHarmonic {k, cos_x0} (from CosineHarmonic::new); Torsion {k, delta, n}.
// A TableComputation (computations.rs:70-175) wraps any pair potential: the
// tables hold energy/force at r = i * delta, and `tabulated` switches
// energy/force to the table while the tails stay with the wrapped potential.
struct Potential {
    int kind = POT_NULL;
    double p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0;
    bool tabulated = false;
    double t_delta = 0, t_max = 0;
    std::vector<double> t_energy, t_force;

    // functions.rs:79-89 (LJ), :135-145 (Harmonic), :31-38 (Null),
    // :198-207 (CosineHarmonic), :244-256 (Torsion), :286-300 (Buckingham),
    // :353-367 (BornMayerHuggins), :414-424 (Morse), :474-482 (Gaussian),
    // :552-566 (Mie)
    double direct_energy(double r) const {
        switch (kind) {
            case POT_LJ: {
                double x = p0 / r;
                double x2 = x * x;
                double x4 = x2 * x2;
                double s6 = x2 * x4;  // powi(x, 6) by binary exponentiation
                return 4.0 * p1 * (s6 * s6 - s6);
            }
            case POT_HARMONIC: {
                double dx = r - p1;
                return 0.5 * p0 * dx * dx;
            }
            case POT_BUCKINGHAM: {
                double r3 = r * r * r;
                double r6 = r3 * r3;
                double ex = std::exp(-r / p2);
                return p0 * ex - p1 / r6;
            }
            case POT_BORN: {
                double r2 = r * r;
                double r6 = r2 * r2 * r2;
                double ex = std::exp((p3 - r) / p4);
                return p0 * ex - p1 / r6 + p2 / (r6 * r2);
            }
            case POT_MORSE: {
                double rc = 1.0 - std::exp((p1 - r) * p0);
                return p2 * rc * rc;
            }
            case POT_GAUSSIAN: return -p0 * std::exp(-p1 * r * r);
            case POT_MIE: {
                double sigma_r = p0 / r;
                double rep = std::pow(sigma_r, p1);
                double att = std::pow(sigma_r, p2);
                return p3 * (rep - att);
            }
            case POT_COSINE_HARMONIC: {
                double dr = std::cos(r) - p1;
                return 0.5 * p0 * dr * dr;
            }
            case POT_TORSION: {
                double c = std::cos(p2 * r - p1);
                return p0 * (1.0 + c);
            }
        }
        return 0.0;
    }
    double direct_force(double r) const {
        switch (kind) {
            case POT_LJ: {
                double x = p0 / r;
                double x2 = x * x;
                double x4 = x2 * x2;
                double s6 = x2 * x4;
                return -24.0 * p1 * (s6 - 2.0 * (s6 * s6)) / r;
            }
            case POT_HARMONIC: return p0 * (p1 - r);
            case POT_BUCKINGHAM: {
                double r3 = r * r * r;
                double r7 = r3 * r3 * r;
                double ex = std::exp(-r / p2);
                return p0 / p2 * ex - 6.0 * p1 / r7;
            }
            case POT_BORN: {
                double r2 = r * r;
                double r7 = r2 * r2 * r2 * r;
                double ex = std::exp((p3 - r) / p4);
                return p0 / p4 * ex - 6.0 * p1 / r7 + 8.0 * p2 / (r7 * r2);
            }
            case POT_MORSE: {
                double ex = std::exp((p1 - r) * p0);
                return 2.0 * p2 * (1.0 - ex * ex) * p0;
            }
            case POT_GAUSSIAN: return 2.0 * p1 * r * direct_energy(r);
            case POT_MIE: {
                double sigma_r = p0 / r;
                double rep = std::pow(sigma_r, p1);
                double att = std::pow(sigma_r, p2);
                return p3 * (p1 * rep - p2 * att) / r;
            }
            case POT_COSINE_HARMONIC: return p0 * (std::cos(r) - p1) * std::sin(r);
            case POT_TORSION: return p0 * p2 * std::sin(p2 * r - p1);
        }
        return 0.0;
    }
    // computations.rs:122-146: linear interpolation in the bin of r; zero
    // from the last bin on.  `floor(r / delta) as usize` saturates at 0.
    static double table_lookup(const std::vector<double>& t, double delta, double r) {
        double q = std::floor(r / delta);
        size_t bin = q > 0.0 ? size_t(q) : 0;  // NaN and negatives -> 0
        if (q >= 1e18) return 0.0;
        if (bin < t.size() - 1) {
            double dx = r - double(bin) * delta;
            double slope = (t[bin + 1] - t[bin]) / delta;
            return t[bin] + dx * slope;
        }
        return 0.0;
    }
    double energy(double r) const { return tabulated ? table_lookup(t_energy, t_delta, r) : direct_energy(r); }
    double force(double r) const { return tabulated ? table_lookup(t_force, t_delta, r) : direct_force(r); }
    // computations.rs:96-117 (TableComputation::new)
    void make_table(size_t size, double max) {
        t_delta = max / double(size);
        t_max = max;
        t_energy.resize(size);
        t_force.resize(size);
        for (size_t i = 0; i < size; i++) {
            double r = double(i) * t_delta;
            t_energy[i] = direct_energy(r);
            t_force[i] = direct_force(r);
        }
        tabulated = true;
    }
    // functions.rs:92-106 (LJ), :302-327 (Buckingham, with its `+ 8.0`),
    // :369-395 (BornMayerHuggins), :484-499 (Gaussian), :568-592 (Mie);
    // zero for the others (Null, Harmonic, Morse).  A table keeps the tails
    // of the potential it wraps (computations.rs:148-175).
    double tail_energy(double rc) const {
        switch (kind) {
            case POT_LJ: {
                double s3 = p0 * p0 * p0;
                double rc3 = rc * rc * rc;
                double s9 = s3 * s3 * s3;
                double rc9 = rc3 * rc3 * rc3;
                return 4.0 / 3.0 * p1 * s3 * (1.0 / 3.0 * s9 / rc9 - s3 / rc3);
            }
            case POT_BUCKINGHAM: {
                double rc2 = rc * rc;
                double rc3 = rc2 * rc;
                double ex = std::exp(-rc / p2);
                double factor = rc2 - 2.0 * rc * p2 + 2.0 * p2 * p2;
                return p0 * p2 * ex * factor - p1 / (3.0 * rc3);
            }
            case POT_BORN: {
                double rc2 = rc * rc;
                double rc3 = rc2 * rc;
                double ex = std::exp((p3 - rc) / p4);
                double factor = rc2 - 2.0 * rc * p4 + 2.0 * p4 * p4;
                return p0 * p4 * ex * factor - p1 / (3.0 * rc3) + p2 / (5.0 * rc2 * rc3);
            }
            case POT_GAUSSIAN:
                return direct_energy(rc) * rc / (2.0 * p1) -
                       p0 * std::sqrt(PI) * std::erfc(std::sqrt(p1) * rc) / (4.0 * std::pow(p1, 3.0 / 2.0));
            case POT_MIE: {
                if (p2 <= 3.0) return 0.0;
                double sigma_rc = p0 / rc;
                double n_3 = p1 - 3.0;
                double m_3 = p2 - 3.0;
                double rep = std::pow(sigma_rc, n_3);
                double att = std::pow(sigma_rc, m_3);
                return p3 * (p0 * p0 * p0) * (rep / n_3 - att / m_3);
            }
        }
        return 0.0;
    }
    double tail_virial(double rc) const {
        switch (kind) {
            case POT_LJ: {
                double s3 = p0 * p0 * p0;
                double rc3 = rc * rc * rc;
                double s9 = s3 * s3 * s3;
                double rc9 = rc3 * rc3 * rc3;
                return 8.0 * p1 * s3 * (2.0 / 3.0 * s9 / rc9 - s3 / rc3);
            }
            case POT_BUCKINGHAM: {
                double rc2 = rc * rc;
                double rc3 = rc2 * rc;
                double ex = std::exp(-rc / p2);
                double factor = rc3 + 3.0 * rc2 * p2 + 6.0 * rc * p2 * p2 + 6.0 * p2 * p2 * p2;
                return p0 * ex * factor - 20.0 * p1 / rc3 + 8.0;
            }
            case POT_BORN: {
                double rc2 = rc * rc;
                double rc3 = rc2 * rc;
                double ex = std::exp((p3 - rc) / p4);
                double factor = rc3 + 3.0 * rc2 * p4 + 6.0 * rc * p4 * p4 + 6.0 * p4 * p4 * p4;
                return p0 * ex * factor - 20.0 * p1 / rc3 + 8.0 * p2 / (5.0 * rc2 * rc3);
            }
            case POT_GAUSSIAN:
                return 3.0 * std::sqrt(PI) * p0 * std::erfc(std::sqrt(p1) * rc) / (4.0 * std::pow(p1, 3.0 / 2.0)) -
                       direct_energy(rc) * rc * (2.0 * p1 * rc * rc + 3.0) / (2.0 * p1);
            case POT_MIE: {
                if (p2 <= 3.0) return 0.0;
                double sigma_rc = p0 / rc;
                double n_3 = p1 - 3.0;
                double m_3 = p2 - 3.0;
                double rep = std::pow(sigma_rc, n_3);
                double att = std::pow(sigma_rc, m_3);
                return p3 * (p0 * p0 * p0) * (rep * p1 / n_3 - att * p2 / m_3);
            }
        }
        return 0.0;
    }
};

// Parameters as the C ABI passes them -> the reference's fields.  Returns
// false for invalid parameters, where the reference's constructors assert
// (Gaussian::new b > 0, functions.rs:465; Mie::new m < n, :544).
bool set_potential(Potential& p, int kind, const double* a) {
    p = Potential{};
    p.kind = kind;
    switch (kind) {
        case POT_NULL: return true;
        case POT_LJ: case POT_HARMONIC: case POT_GAUSSIAN:
            p.p0 = a[0];
            p.p1 = a[1];
            return kind != POT_GAUSSIAN || p.p1 > 0.0;
        case POT_BUCKINGHAM: case POT_MORSE:
            p.p0 = a[0];
            p.p1 = a[1];
            p.p2 = a[2];
            return true;
        case POT_BORN:
            p.p0 = a[0];
            p.p1 = a[1];
            p.p2 = a[2];
            p.p3 = a[3];
            p.p4 = a[4];
            return true;
        case POT_MIE: {  // a = {sigma, epsilon, n, m}; functions.rs:543-552
            double sigma = a[0], eps = a[1], n = a[2], m = a[3];
            if (!(m < n)) return false;
            p.p0 = sigma;
            p.p1 = n;
            p.p2 = m;
            p.p3 = n / (n - m) * std::pow(n / m, m / (n - m)) * eps;
            return true;
        }
        case POT_COSINE_HARMONIC:  // a = {k, x0}; functions.rs:189-196
            p.p0 = a[0];
            p.p1 = std::cos(a[1]);
            return true;
        case POT_TORSION:  // a = {k, delta, n}
            p.p0 = a[0];
            p.p1 = a[1];
            p.p2 = double(uint64_t(a[2]));
            return a[2] >= 0.0 && a[2] == std::floor(a[2]);
    }
    return false;
}

// restrictions.rs:13-33
enum Restriction {
    R_NONE = 0,
    R_INTRA = 1,
    R_INTER = 2,
    R_EXCLUDE12 = 3,
    R_EXCLUDE13 = 4,
    R_EXCLUDE14 = 5,
    R_SCALE14 = 6
};
// restrictions.rs:37-50
enum BondPath { BP_NONE = 0, BP_SAME = 1, BP_ONE = 2, BP_TWO = 3, BP_THREE = 4, BP_FAR = 5 };

struct RInfo {
    bool excluded;
    double scaling;
};
// restrictions.rs:85-116
RInfo information(int restriction, double scale14, int path) {
    bool same = path != BP_NONE;
    bool excluded = false;
    switch (restriction) {
        case R_NONE: excluded = false; break;
        case R_INTER: excluded = same; break;
        case R_INTRA: excluded = !same; break;
        case R_EXCLUDE12: excluded = path == BP_ONE; break;
        case R_EXCLUDE13:
        case R_SCALE14: excluded = path == BP_ONE || path == BP_TWO; break;
        case R_EXCLUDE14: excluded = path == BP_ONE || path == BP_TWO || path == BP_THREE; break;
    }
    double scaling = 1.0;
    if (restriction == R_SCALE14 && path == BP_THREE) scaling = scale14;
    return {excluded, scaling};
}

// pairs.rs:27-297
struct PairInteraction {
    Potential pot;
    double cutoff = 0;
    bool shifted = false;
    double shift = 0;
    bool tail = false;
    int restriction = R_NONE;
    double scale14 = 1.0;

    double energy(double r) const {
        if (r >= cutoff) return 0.0;
        double e = pot.energy(r);
        return shifted ? e - shift : e;
    }
    double force(double r) const {
        if (r >= cutoff) return 0.0;
        return pot.force(r);
    }
    // energy/mod.rs:137-142 default virial, gated by pairs.rs:237-243
    M3 virial(V3 d) const {
        if (norm(d) >= cutoff) return M3{};
        double fact = pot.force(norm(d));
        V3 rn = d / norm(d);
        V3 f = fact * rn;
        return tensorial(f, d);
    }
    double tail_energy() const { return tail ? pot.tail_energy(cutoff) : 0.0; }
    double tail_virial_scalar() const { return tail ? pot.tail_virial(cutoff) : 0.0; }
};

// --- Ewald parameters and factors (ewald.rs:80-183, 278-378) ----------------
struct EwaldFactor {
    int ikx, iky, ikz;
    double energy;
    V3 field;
    M3 virial;
};

struct Ewald {
    double alpha = 0, rc = 0, kmax2 = 0;
    int kmax = 0;
    int restriction = R_NONE;
    double scale14 = 1.0;
    std::vector<EwaldFactor> factors;
    double prev_a = -1, prev_b = -1, prev_c = -1;
    // cached rho(k) for MC (ewald.rs:222) and the staged delta
    std::vector<Cx> rho;
    std::vector<Cx> staged_delta;
    bool has_staged = false;
};

struct Wolf {
    double alpha = 0, cutoff = 0, energy_constant = 0, force_constant = 0;
    int restriction = R_NONE;
    double scale14 = 1.0;
};

// --- the system --------------------------------------------------------------
struct System {
    size_t n = 0;
    double L[3] = {0, 0, 0};  // orthorhombic lengths; 0 = infinite
    std::vector<V3> pos, vel;
    std::vector<double> mass, charge;
    std::vector<uint32_t> kind;
    uint32_t nkinds = 0;
    // molecules: contiguous ranges, molecule id per atom
    std::vector<uint32_t> molid;
    std::vector<size_t> mol_start;  // nmol + 1
    std::vector<std::pair<size_t, size_t>> bonds;
    std::vector<std::tuple<size_t, size_t, size_t>> angles;
    std::vector<std::tuple<size_t, size_t, size_t, size_t>> dihedrals;
    // per-molecule bond distance flags, keyed by (i,j) global indices, i != j
    std::map<std::pair<size_t, size_t>, int> bond_flags;

    std::map<std::pair<uint32_t, uint32_t>, PairInteraction> pairs;
    std::map<std::pair<uint32_t, uint32_t>, Potential> bond_pots;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, Potential> angle_pots;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t>, Potential> dihedral_pots;

    int coulomb = 0;  // 0 none, 1 ewald, 2 wolf
    Ewald ewald;
    Wolf wolf;
    int nthreads = 1;

    // MD state
    std::vector<V3> acc;

    // MC energy cache (cache.rs:22-48)
    std::vector<double> pair_cache;  // n*n
    double c_pairs = 0, c_tail = 0, c_bonds = 0, c_angles = 0, c_coulomb = 0;
    bool cache_ready = false;
    bool staged = false;
    size_t staged_mol = 0;
    std::vector<V3> staged_pos;
    std::vector<double> staged_new_pairs;  // per (moved atom, j)
    double staged_pairs_delta = 0, staged_coulomb_delta = 0;

    bool infinite() const { return L[0] == 0.0; }

    // cells.rs:284-300, orthorhombic branch
    void vector_image(V3& d) const {
        if (infinite()) return;
        d.x -= std::round(d.x / L[0]) * L[0];
        d.y -= std::round(d.y / L[1]) * L[1];
        d.z -= std::round(d.z / L[2]) * L[2];
    }
    // configuration.rs:399-403
    V3 nearest_image(size_t i, size_t j) const {
        V3 r = pos[i] - pos[j];
        vector_image(r);
        return r;
    }
    // cells.rs:316-320
    double cell_distance(V3 u, V3 v) const {
        V3 d = v - u;
        vector_image(d);
        return norm(d);
    }
    double distance(size_t i, size_t j) const { return cell_distance(pos[i], pos[j]); }
    double volume() const { return infinite() ? 0.0 : L[0] * L[1] * L[2]; }

    // configuration.rs:124-144
    int bond_path(size_t i, size_t j) const {
        if (molid[i] != molid[j]) return BP_NONE;
        if (i == j) return BP_SAME;
        auto it = bond_flags.find({i, j});
        int flags = it == bond_flags.end() ? 0 : it->second;
        if (flags & 1) return BP_ONE;
        if (flags & 2) return BP_TWO;
        if (flags & 4) return BP_THREE;
        return BP_FAR;
    }

    const PairInteraction* pair_potential(size_t i, size_t j) const {
        uint32_t a = kind[i], b = kind[j];
        if (a > b) std::swap(a, b);
        auto it = pairs.find({a, b});
        return it == pairs.end() ? nullptr : &it->second;
    }
    const Potential* bond_potential(size_t i, size_t j) const {
        uint32_t a = kind[i], b = kind[j];
        if (a > b) std::swap(a, b);
        auto it = bond_pots.find({a, b});
        return it == bond_pots.end() ? nullptr : &it->second;
    }
    const Potential* angle_potential(size_t i, size_t j, size_t k) const {
        uint32_t a = kind[i], b = kind[j], c = kind[k];
        if (a > c) std::swap(a, c);
        auto it = angle_pots.find({a, b, c});
        return it == angle_pots.end() ? nullptr : &it->second;
    }
    // interactions.rs:44-56 normalize_dihedral
    static std::tuple<uint32_t, uint32_t, uint32_t, uint32_t> dihedral_key(uint32_t i, uint32_t j, uint32_t k,
                                                                          uint32_t m) {
        uint32_t ij = std::max(i, j), km = std::max(k, m);
        if (ij == km) {
            if (std::min(i, j) < std::min(k, m)) return {i, j, k, m};
            return {m, k, j, i};
        }
        if (ij < km) return {i, j, k, m};
        return {m, k, j, i};
    }
    const Potential* dihedral_potential(size_t i, size_t j, size_t k, size_t m) const {
        auto it = dihedral_pots.find(dihedral_key(kind[i], kind[j], kind[k], kind[m]));
        return it == dihedral_pots.end() ? nullptr : &it->second;
    }
};

// Static interleaved partition of the i loop over threads, as rayon splits
// the (0..natoms) range; each worker owns a thread-local buffer.
template <class F>
void parallel_for_i(int nthreads, size_t n, F&& f) {
    if (nthreads <= 1 || n < 64) {
        f(0, size_t(0), n, size_t(1));
        return;
    }
    std::vector<std::thread> threads;
    for (int t = 0; t < nthreads; t++) {
        threads.emplace_back([&, t]() { f(t, size_t(t), n, size_t(nthreads)); });
    }
    for (auto& th : threads) th.join();
}

// --- pair terms ---------------------------------------------------------------
// compute.rs:33-56
void pair_forces(const System& s, std::vector<V3>& out) {
    size_t n = s.n;
    int nt = std::max(1, s.nthreads);
    std::vector<std::vector<V3>> local(nt, std::vector<V3>(n));
    parallel_for_i(nt, n, [&](int t, size_t i0, size_t n_, size_t step) {
        auto& f = local[t];
        for (size_t i = i0; i < n_; i += step) {
            V3 fi;
            for (size_t j = i + 1; j < n; j++) {
                int path = s.bond_path(i, j);
                V3 d = s.nearest_image(i, j);
                V3 dn = d / norm(d);
                double r = norm(d);
                const PairInteraction* p = s.pair_potential(i, j);
                if (p) {
                    RInfo info = information(p->restriction, p->scale14, path);
                    if (!info.excluded) {
                        V3 force = (info.scaling * p->force(r)) * dn;
                        fi += force;
                        f[j] -= force;
                    }
                }
            }
            f[i] += fi;
        }
    });
    for (int t = 0; t < nt; t++)
        for (size_t i = 0; i < n; i++) out[i] += local[t][i];
}

// energy.rs:32-59
double pair_energy_of(const System& s, int path, double r, size_t i, size_t j) {
    const PairInteraction* p = s.pair_potential(i, j);
    if (!p) return 0.0;
    RInfo info = information(p->restriction, p->scale14, path);
    if (info.excluded) return 0.0;
    return info.scaling * p->energy(r);
}

double pairs_energy(const System& s) {
    size_t n = s.n;
    int nt = std::max(1, s.nthreads);
    std::vector<double> partial(nt, 0.0);
    parallel_for_i(nt, n, [&](int t, size_t i0, size_t n_, size_t step) {
        double acc = 0.0;
        for (size_t i = i0; i < n_; i += step) {
            double local = 0.0;
            for (size_t j = i + 1; j < n; j++) {
                double r = norm(s.nearest_image(i, j));
                int path = s.bond_path(i, j);
                local += pair_energy_of(s, path, r, i, j);
            }
            acc += local;
        }
        partial[t] = acc;
    });
    double e = 0.0;
    for (double p : partial) e += p;
    return e;
}

// composition.rs:123-128 (kinds with non-zero count, in kind order)
std::vector<std::pair<uint32_t, double>> composition(const System& s) {
    std::vector<size_t> count(s.nkinds, 0);
    for (size_t i = 0; i < s.n; i++) count[s.kind[i]]++;
    std::vector<std::pair<uint32_t, double>> out;
    for (uint32_t k = 0; k < s.nkinds; k++)
        if (count[k] != 0) out.push_back({k, double(count[k])});
    return out;
}

// energy.rs:63-79
double pairs_tail(const System& s) {
    if (s.infinite()) return 0.0;
    double e = 0.0;
    double volume = s.volume();
    auto comp = composition(s);
    for (auto& [ki, ni] : comp) {
        for (auto& [kj, nj] : comp) {
            double two_pi_density = 2.0 * PI * ni * nj / volume;
            uint32_t a = std::min(ki, kj), b = std::max(ki, kj);
            auto it = s.pairs.find({a, b});
            if (it != s.pairs.end()) e += two_pi_density * it->second.tail_energy();
        }
    }
    return e;
}

// compute.rs:199-215
M3 pairs_virial(const System& s) {
    size_t n = s.n;
    int nt = std::max(1, s.nthreads);
    std::vector<M3> partial(nt);
    parallel_for_i(nt, n, [&](int t, size_t i0, size_t n_, size_t step) {
        M3 acc;
        for (size_t i = i0; i < n_; i += step) {
            M3 local;
            for (size_t j = i + 1; j < n; j++) {
                int path = s.bond_path(i, j);
                const PairInteraction* p = s.pair_potential(i, j);
                if (p) {
                    RInfo info = information(p->restriction, p->scale14, path);
                    if (!info.excluded) {
                        V3 d = s.nearest_image(i, j);
                        local += scale(info.scaling, p->virial(d));
                    }
                }
            }
            acc += local;
        }
        partial[t] = acc;
    });
    M3 w;
    for (auto& p : partial) w += p;
    // compute.rs:218-228 tail: two_pi_density * tail_virial * I/3
    double volume = s.volume();
    auto comp = composition(s);
    for (auto& [ki, ni] : comp) {
        for (auto& [kj, nj] : comp) {
            double two_pi_density = 2.0 * PI * ni * nj / volume;
            uint32_t a = std::min(ki, kj), b = std::max(ki, kj);
            auto it = s.pairs.find({a, b});
            if (it != s.pairs.end()) {
                double tv = it->second.tail_virial_scalar();
                double diag = tv * (1.0 / 3.0);
                for (int k = 0; k < 3; k++) w.m[k][k] += two_pi_density * diag;
                // off-diagonal terms are two_pi_density * (tv * 0) == 0
            }
        }
    }
    return w;
}

// --- bonded terms (compute.rs:60-85, energy.rs:88-117, cells.rs:335-359) ----------
void angle_and_derivatives(const System& s, V3 r1, V3 r2, V3 r3, double& theta, V3& d1, V3& d2,
                           V3& d3) {
    V3 r12 = r1 - r2;
    s.vector_image(r12);
    V3 r23 = r3 - r2;
    s.vector_image(r23);
    double r12_norm = norm(r12);
    double r23_norm = norm(r23);
    V3 r12n = r12 / r12_norm;
    V3 r23n = r23 / r23_norm;
    double cos = dot(r12n, r23n);
    double sin_inv = 1.0 / std::sqrt(1.0 - cos * cos);
    d1 = (sin_inv * (cos * r12n - r23n)) / r12_norm;
    d3 = (sin_inv * (cos * r23n - r12n)) / r23_norm;
    V3 sum = d1 + d3;
    d2 = V3{-sum.x, -sum.y, -sum.z};
    theta = std::acos(cos);
}

double angle_value(const System& s, size_t i, size_t j, size_t k) {
    V3 r12 = s.pos[i] - s.pos[j];
    s.vector_image(r12);
    V3 r23 = s.pos[k] - s.pos[j];
    s.vector_image(r23);
    return std::acos(dot(r12, r23) / (norm(r12) * norm(r23)));
}

inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// cells.rs:364-411
void dihedral_and_derivatives(const System& s, V3 r1, V3 r2, V3 r3, V3 r4, double& phi, V3& d1, V3& d2,
                              V3& d3, V3& d4) {
    V3 r12 = r2 - r1;
    s.vector_image(r12);
    V3 r23 = r3 - r2;
    s.vector_image(r23);
    V3 r34 = r4 - r3;
    s.vector_image(r34);
    V3 u = cross(r12, r23);
    V3 v = cross(r23, r34);
    double u_norm2 = norm2(u);
    double v_norm2 = norm2(v);
    double r23_norm2 = norm2(r23);
    double r23_norm = std::sqrt(r23_norm2);
    d1 = (-r23_norm / u_norm2) * u;
    d4 = (r23_norm / v_norm2) * v;
    double r23_r34 = dot(r23, r34);
    double r12_r23 = dot(r12, r23);
    d2 = ((-r12_r23 / r23_norm2 - 1.0) * d1) + ((r23_r34 / r23_norm2) * d4);
    d3 = ((-r23_r34 / r23_norm2 - 1.0) * d4) + ((r12_r23 / r23_norm2) * d1);
    phi = std::atan2(dot(r23_norm * v, r12), dot(u, v));
}

double dihedral_value(const System& s, size_t i, size_t j, size_t k, size_t m) {
    V3 r12 = s.pos[j] - s.pos[i];
    s.vector_image(r12);
    V3 r23 = s.pos[k] - s.pos[j];
    s.vector_image(r23);
    V3 r34 = s.pos[m] - s.pos[k];
    s.vector_image(r34);
    V3 u = cross(r12, r23);
    V3 v = cross(r23, r34);
    return std::atan2(dot(norm(r23) * v, r12), dot(u, v));
}

double dihedrals_energy(const System& s) {
    double e = 0.0;
    for (auto& [i, j, k, m] : s.dihedrals) {
        double phi = dihedral_value(s, i, j, k, m);
        const Potential* p = s.dihedral_potential(i, j, k, m);
        if (p) e += p->energy(phi);
    }
    return e;
}

void bonded_forces(const System& s, std::vector<V3>& f) {
    for (auto& [i, j] : s.bonds) {
        V3 d = s.nearest_image(i, j);
        V3 dn = d / norm(d);
        double r = norm(d);
        const Potential* p = s.bond_potential(i, j);
        if (p) {
            V3 force = p->force(r) * dn;
            f[i] += force;
            f[j] -= force;
        }
    }
    for (auto& [i, j, k] : s.angles) {
        double theta;
        V3 d1, d2, d3;
        angle_and_derivatives(s, s.pos[i], s.pos[j], s.pos[k], theta, d1, d2, d3);
        const Potential* p = s.angle_potential(i, j, k);
        if (p) {
            double force = p->force(theta);
            f[i] += force * d1;
            f[j] += force * d2;
            f[k] += force * d3;
        }
    }
    for (auto& [i, j, k, m] : s.dihedrals) {
        double phi;
        V3 d1, d2, d3, d4;
        dihedral_and_derivatives(s, s.pos[i], s.pos[j], s.pos[k], s.pos[m], phi, d1, d2, d3, d4);
        const Potential* p = s.dihedral_potential(i, j, k, m);
        if (p) {
            double force = p->force(phi);
            f[i] += force * d1;
            f[j] += force * d2;
            f[k] += force * d3;
            f[m] += force * d4;
        }
    }
}

double bonds_energy(const System& s) {
    double e = 0.0;
    for (auto& [i, j] : s.bonds) {
        double r = norm(s.nearest_image(i, j));
        const Potential* p = s.bond_potential(i, j);
        if (p) e += p->energy(r);
    }
    return e;
}

double angles_energy(const System& s) {
    double e = 0.0;
    for (auto& [i, j, k] : s.angles) {
        double theta = angle_value(s, i, j, k);
        const Potential* p = s.angle_potential(i, j, k);
        if (p) e += p->energy(theta);
    }
    return e;
}

M3 bonds_virial(const System& s) {
    M3 w;
    for (auto& [i, j] : s.bonds) {
        V3 r = s.nearest_image(i, j);
        const Potential* p = s.bond_potential(i, j);
        if (p) {
            // default PairPotential-like virial: (f(r) r/|r|) x r
            double fact = p->force(norm(r));
            V3 rn = r / norm(r);
            w += tensorial(fact * rn, r);
        }
    }
    return w;
}

// --- Ewald (ewald.rs) --------------------------------------------------------------
void ewald_prepare(const System& s, Ewald& ew) {
    if (ew.prev_a == s.L[0] && ew.prev_b == s.L[1] && ew.prev_c == s.L[2]) return;
    ew.prev_a = s.L[0];
    ew.prev_b = s.L[1];
    ew.prev_c = s.L[2];
    // k_vector(index) = 2 pi * inv * index, inv from Matrix3::inverse
    // (matrix.rs:212-226): inv[k][k] = (product of the two others) / det.
    double det = s.L[0] * (s.L[1] * s.L[2] - 0.0 * 0.0) - 0.0 + 0.0;
    double inv_det = 1.0 / det;
    double inv[3] = {(s.L[1] * s.L[2] - 0.0 * 0.0) * inv_det, (s.L[0] * s.L[2] - 0.0 * 0.0) * inv_det,
                     (s.L[0] * s.L[1] - 0.0 * 0.0) * inv_det};
    auto kvec = [&](double a, double b, double c) {
        return V3{2.0 * PI * inv[0] * a, 2.0 * PI * inv[1] * b, 2.0 * PI * inv[2] * c};
    };
    V3 k111 = kvec(1.0, 1.0, 1.0);
    double mx = std::max(k111.x, std::max(k111.y, k111.z)) * double(ew.kmax);
    ew.kmax2 = 1.0001 * mx * mx;

    double alpha_sq_inv_fourth = 0.25 / (ew.alpha * ew.alpha);
    double four_pi_v = 4.0 * PI / s.volume();
    ew.factors.clear();
    auto push = [&](int ikx, int iky, int ikz) {
        V3 k = kvec(ikx, iky, ikz);
        double k2 = norm2(k);
        if (k2 > ew.kmax2) return;
        EwaldFactor f;
        f.ikx = ikx;
        f.iky = iky;
        f.ikz = ikz;
        f.energy = four_pi_v * std::exp(-k2 * alpha_sq_inv_fourth) / k2;
        f.field = (2.0 * f.energy) * k;
        double vf = -2.0 * (1.0 / k2 + alpha_sq_inv_fourth);
        M3 kk = tensorial(k, k);
        M3 v;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) v.m[a][b] = (a == b ? 1.0 : 0.0) + vf * kk.m[a][b];
        f.virial = scale(f.energy, v);
        ew.factors.push_back(f);
    };
    int kmax = ew.kmax;
    for (int ikx = 1; ikx < kmax; ikx++)
        for (int iky = -kmax; iky < kmax; iky++)
            for (int ikz = -kmax; ikz < kmax; ikz++) push(ikx, iky, ikz);
    for (int iky = 1; iky < kmax; iky++)
        for (int ikz = -kmax; ikz < kmax; ikz++) push(0, iky, ikz);
    for (int ikz = 1; ikz < kmax; ikz++) push(0, 0, ikz);
}

// ewald.rs:387-401
double ewald_real_energy_pair(const Ewald& ew, RInfo info, double qiqj, double r) {
    if (r > ew.rc) return 0.0;
    if (info.excluded) return -qiqj / FOUR_PI_EPSILON_0 * std::erf(ew.alpha * r) / r;
    return qiqj / FOUR_PI_EPSILON_0 * std::erfc(ew.alpha * r) / r;
}
// ewald.rs:408-428
double ewald_real_force_pair(const Ewald& ew, RInfo info, double qiqj, double r) {
    if (r > ew.rc) return 0.0;
    if (info.excluded)
        return qiqj / (FOUR_PI_EPSILON_0 * r * r) *
               (ew.alpha * FRAC_2_SQRT_PI * std::exp(-ew.alpha * ew.alpha * r * r) - std::erf(ew.alpha * r) / r);
    return qiqj / (FOUR_PI_EPSILON_0 * r * r) *
           (ew.alpha * FRAC_2_SQRT_PI * std::exp(-ew.alpha * ew.alpha * r * r) + std::erfc(ew.alpha * r) / r);
}

double ewald_real_energy(const System& s, const Ewald& ew) {
    size_t n = s.n;
    int nt = std::max(1, s.nthreads);
    std::vector<double> partial(nt, 0.0);
    parallel_for_i(nt, n, [&](int t, size_t i0, size_t n_, size_t step) {
        double acc = 0.0;
        for (size_t i = i0; i < n_; i += step) {
            double local = 0.0;
            double qi = s.charge[i];
            if (qi == 0.0) continue;
            for (size_t j = i + 1; j < n; j++) {
                double qj = s.charge[j];
                if (qj == 0.0) continue;
                RInfo info = information(ew.restriction, ew.scale14, s.bond_path(i, j));
                double r = s.distance(i, j);
                local += ewald_real_energy_pair(ew, info, qi * qj, r);
            }
            acc += local;
        }
        partial[t] = acc;
    });
    double e = 0.0;
    for (double p : partial) e += p;
    return e;
}

void ewald_real_forces(const System& s, const Ewald& ew, std::vector<V3>& out) {
    size_t n = s.n;
    int nt = std::max(1, s.nthreads);
    std::vector<std::vector<V3>> local(nt, std::vector<V3>(n));
    parallel_for_i(nt, n, [&](int t, size_t i0, size_t n_, size_t step) {
        auto& f = local[t];
        for (size_t i = i0; i < n_; i += step) {
            V3 fi;
            double qi = s.charge[i];
            if (qi == 0.0) continue;
            for (size_t j = i + 1; j < n; j++) {
                double qj = s.charge[j];
                if (qj == 0.0) continue;
                RInfo info = information(ew.restriction, ew.scale14, s.bond_path(i, j));
                V3 rij = s.nearest_image(i, j);
                V3 force = ewald_real_force_pair(ew, info, qi * qj, norm(rij)) * rij;
                fi += force;
                f[j] -= force;
            }
            f[i] += fi;
        }
    });
    for (int t = 0; t < nt; t++)
        for (size_t i = 0; i < n; i++) out[i] += local[t][i];
}

M3 ewald_real_virial(const System& s, const Ewald& ew) {
    size_t n = s.n;
    int nt = std::max(1, s.nthreads);
    std::vector<M3> partial(nt);
    parallel_for_i(nt, n, [&](int t, size_t i0, size_t n_, size_t step) {
        M3 acc;
        for (size_t i = i0; i < n_; i += step) {
            double qi = s.charge[i];
            if (qi == 0.0) continue;
            M3 local;
            for (size_t j = i + 1; j < n; j++) {
                double qj = s.charge[j];
                if (qj == 0.0) continue;
                RInfo info = information(ew.restriction, ew.scale14, s.bond_path(i, j));
                V3 rij = s.nearest_image(i, j);
                V3 force = ewald_real_force_pair(ew, info, qi * qj, norm(rij)) * rij;
                local += tensorial(force, rij);
            }
            acc += local;
        }
        partial[t] = acc;
    });
    M3 w;
    for (auto& p : partial) w += p;
    return w;
}

// ewald.rs:619-626
double ewald_self_energy(const System& s, const Ewald& ew) {
    double q2 = 0.0;
    for (size_t i = 0; i < s.n; i++) q2 += s.charge[i] * s.charge[i];
    return -ew.alpha / std::sqrt(PI) * q2 / FOUR_PI_EPSILON_0;
}

// ewald.rs:633-674: per-axis phase factors by recursion, then rho(k).
// eikr layout: [(n + kmax) * 3 + axis] * natoms + i
struct EikrTable {
    int kmax = 0;
    size_t n = 0;
    std::vector<Cx> data;
    Cx& at(int k, int axis, size_t i) { return data[(size_t(k + kmax) * 3 + axis) * n + i]; }
    Cx at(int k, int axis, size_t i) const { return data[(size_t(k + kmax) * 3 + axis) * n + i]; }
};

void eik_dot_r(const System& s, const Ewald& ew, const std::vector<V3>& positions, EikrTable& t) {
    size_t n = positions.size();
    t.kmax = ew.kmax;
    t.n = n;
    t.data.assign(size_t(2 * ew.kmax + 1) * 3 * n, Cx{});
    double inv[3] = {(s.L[1] * s.L[2]) / (s.L[0] * (s.L[1] * s.L[2])), 0, 0};
    double det = s.L[0] * (s.L[1] * s.L[2] - 0.0 * 0.0) - 0.0 + 0.0;
    double inv_det = 1.0 / det;
    inv[0] = (s.L[1] * s.L[2] - 0.0 * 0.0) * inv_det;
    inv[1] = (s.L[0] * s.L[2] - 0.0 * 0.0) * inv_det;
    inv[2] = (s.L[0] * s.L[1] - 0.0 * 0.0) * inv_det;
    for (int spatial = 0; spatial < 3; spatial++) {
        V3 kv{0, 0, 0};
        kv[spatial] = 2.0 * PI * inv[spatial] * 1.0;
        for (size_t i = 0; i < n; i++) {
            t.at(0, spatial, i) = Cx{1.0, 0.0};
            t.at(1, spatial, i) = polar(1.0, dot(kv, positions[i]));
            t.at(-1, spatial, i) = conj(t.at(1, spatial, i));
        }
    }
    for (int spatial = 0; spatial < 3; spatial++)
        for (int k = 2; k <= ew.kmax; k++)
            for (size_t i = 0; i < n; i++) {
                t.at(k, spatial, i) = cmul(t.at(k - 1, spatial, i), t.at(1, spatial, i));
                t.at(-k, spatial, i) = conj(t.at(k, spatial, i));
            }
}

void compute_rho(const System& s, const Ewald& ew, const EikrTable& t, std::vector<Cx>& rho) {
    rho.clear();
    for (auto& f : ew.factors) {
        Cx partial;
        for (size_t i = 0; i < s.n; i++) {
            Cx phi = cmul(cmul(t.at(f.ikx, 0, i), t.at(f.iky, 1, i)), t.at(f.ikz, 2, i));
            partial.re += s.charge[i] * phi.re;
            partial.im += s.charge[i] * phi.im;
        }
        rho.push_back(partial);
    }
}

double ewald_k_energy(const System& s, Ewald& ew) {
    EikrTable t;
    eik_dot_r(s, ew, s.pos, t);
    compute_rho(s, ew, t, ew.rho);
    double e = 0.0;
    for (size_t k = 0; k < ew.factors.size(); k++) e += ew.factors[k].energy * cnorm2(ew.rho[k]);
    return e / FOUR_PI_EPSILON_0;
}

void ewald_k_forces(const System& s, Ewald& ew, std::vector<V3>& out) {
    EikrTable t;
    eik_dot_r(s, ew, s.pos, t);
    compute_rho(s, ew, t, ew.rho);
    size_t n = s.n;
    size_t nk = ew.factors.size();
    int nt = std::max(1, s.nthreads);
    // ewald.rs:699-712: parallel over k with thread-local field buffers
    std::vector<std::vector<V3>> local(nt, std::vector<V3>(n));
    auto work = [&](int tid, size_t k0, size_t step) {
        auto& field = local[tid];
        for (size_t k = k0; k < nk; k += step) {
            const EwaldFactor& f = ew.factors[k];
            Cx rc = conj(ew.rho[k]);
            for (size_t i = 0; i < n; i++) {
                Cx e = cmul(cmul(t.at(f.ikx, 0, i), t.at(f.iky, 1, i)), t.at(f.ikz, 2, i));
                Cx partial = cmul(e, rc);
                field[i] += partial.im * f.field;
            }
        }
    };
    if (nt <= 1) {
        work(0, 0, 1);
    } else {
        std::vector<std::thread> th;
        for (int q = 0; q < nt; q++) th.emplace_back(work, q, size_t(q), size_t(nt));
        for (auto& x : th) x.join();
    }
    std::vector<V3> field(n);
    for (int q = 0; q < nt; q++)
        for (size_t i = 0; i < n; i++) field[i] += local[q][i];
    for (size_t i = 0; i < n; i++) out[i] += (s.charge[i] * field[i]) / FOUR_PI_EPSILON_0;
}

M3 ewald_k_virial(const System& s, Ewald& ew) {
    EikrTable t;
    eik_dot_r(s, ew, s.pos, t);
    compute_rho(s, ew, t, ew.rho);
    M3 w;
    for (size_t k = 0; k < ew.factors.size(); k++) w += scale(cnorm2(ew.rho[k]), ew.factors[k].virial);
    for (int a = 0; a < 3; a++)  // ewald.rs:732 divides each entry
        for (int b = 0; b < 3; b++) w.m[a][b] = w.m[a][b] / FOUR_PI_EPSILON_0;
    return w;
}

// --- Wolf (wolf.rs) ----------------------------------------------------------------
void wolf_init(Wolf& w, double cutoff) {
    w.alpha = PI / cutoff;
    w.cutoff = cutoff;
    double alpha_cutoff = w.alpha * cutoff;
    double alpha_cutoff_2 = alpha_cutoff * alpha_cutoff;
    w.energy_constant = std::erfc(alpha_cutoff) / cutoff;
    w.force_constant = std::erfc(alpha_cutoff) / (cutoff * cutoff) +
                       FRAC_2_SQRT_PI * w.alpha * std::exp(-alpha_cutoff_2) / cutoff;
}
double wolf_energy_pair(const Wolf& w, double qiqj, double rij) {
    if (rij > w.cutoff) return 0.0;
    return qiqj * (std::erfc(w.alpha * rij) / rij - w.energy_constant) / FOUR_PI_EPSILON_0;
}
double wolf_energy_self(const Wolf& w, double qi) {
    return qi * qi * 0.5 * (w.energy_constant + w.alpha * FRAC_2_SQRT_PI) / FOUR_PI_EPSILON_0;
}
double wolf_force_pair(const Wolf& w, double qiqj, double rij) {
    if (rij > w.cutoff) return 0.0;
    double rij2 = rij * rij;
    double alpha_rij = w.alpha * rij;
    double exp_alpha_rij = std::exp(-alpha_rij * alpha_rij);
    double factor = std::erfc(alpha_rij) / rij2 + w.alpha * FRAC_2_SQRT_PI * exp_alpha_rij / rij;
    return qiqj * (factor - w.force_constant) / (rij * FOUR_PI_EPSILON_0);
}

double wolf_energy(const System& s, const Wolf& w) {
    size_t n = s.n;
    int nt = std::max(1, s.nthreads);
    std::vector<double> partial(nt, 0.0);
    parallel_for_i(nt, n, [&](int t, size_t i0, size_t n_, size_t step) {
        double acc = 0.0;
        for (size_t i = i0; i < n_; i += step) {
            double e = 0.0;
            double qi = s.charge[i];
            if (qi == 0.0) continue;
            for (size_t j = i + 1; j < n; j++) {
                double qj = s.charge[j];
                if (qj == 0.0) continue;
                RInfo info = information(w.restriction, w.scale14, s.bond_path(i, j));
                if (info.excluded) continue;
                double rij = s.distance(i, j);
                e += info.scaling * wolf_energy_pair(w, qi * qj, rij);
            }
            acc += e - wolf_energy_self(w, qi);
        }
        partial[t] = acc;
    });
    double e = 0.0;
    for (double p : partial) e += p;
    return e;
}

void wolf_forces(const System& s, const Wolf& w, std::vector<V3>& out) {
    size_t n = s.n;
    int nt = std::max(1, s.nthreads);
    std::vector<std::vector<V3>> local(nt, std::vector<V3>(n));
    parallel_for_i(nt, n, [&](int t, size_t i0, size_t n_, size_t step) {
        auto& f = local[t];
        for (size_t i = i0; i < n_; i += step) {
            V3 fi;
            double qi = s.charge[i];
            if (qi == 0.0) continue;
            for (size_t j = i + 1; j < n; j++) {
                double qj = s.charge[j];
                if (qj == 0.0) continue;
                RInfo info = information(w.restriction, w.scale14, s.bond_path(i, j));
                if (info.excluded) continue;
                V3 rij = s.nearest_image(i, j);
                V3 force = (info.scaling * wolf_force_pair(w, qi * qj, norm(rij))) * rij;
                fi += force;
                f[j] -= force;
            }
            f[i] += fi;
        }
    });
    for (int t = 0; t < nt; t++)
        for (size_t i = 0; i < n; i++) out[i] += local[t][i];
}

M3 wolf_virial(const System& s, const Wolf& w) {
    size_t n = s.n;
    int nt = std::max(1, s.nthreads);
    std::vector<M3> partial(nt);
    parallel_for_i(nt, n, [&](int t, size_t i0, size_t n_, size_t step) {
        M3 acc;
        for (size_t i = i0; i < n_; i += step) {
            double qi = s.charge[i];
            if (qi == 0.0) continue;
            M3 local;
            for (size_t j = i + 1; j < n; j++) {
                double qj = s.charge[j];
                if (qj == 0.0) continue;
                RInfo info = information(w.restriction, w.scale14, s.bond_path(i, j));
                if (info.excluded) continue;
                V3 rij = s.nearest_image(i, j);
                V3 force = (info.scaling * wolf_force_pair(w, qi * qj, norm(rij))) * rij;
                local += tensorial(force, rij);
            }
            acc += local;
        }
        partial[t] = acc;
    });
    M3 v;
    for (auto& p : partial) v += p;
    return v;
}

// --- System-level compute (compute.rs, energy.rs) --------------------------------
void all_forces(System& s, std::vector<V3>& f) {
    f.assign(s.n, V3{});
    pair_forces(s, f);
    bonded_forces(s, f);
    if (s.coulomb == 1) {
        ewald_prepare(s, s.ewald);
        ewald_real_forces(s, s.ewald, f);
        ewald_k_forces(s, s.ewald, f);
    } else if (s.coulomb == 2) {
        wolf_forces(s, s.wolf, f);
    }
}

double coulomb_energy(System& s) {
    if (s.coulomb == 1) {
        ewald_prepare(s, s.ewald);
        double real = ewald_real_energy(s, s.ewald);
        double self_e = ewald_self_energy(s, s.ewald);
        double k = ewald_k_energy(s, s.ewald);
        return real + self_e + k;
    } else if (s.coulomb == 2) {
        return wolf_energy(s, s.wolf);
    }
    return 0.0;
}

M3 atomic_virial(System& s) {
    M3 w = pairs_virial(s);
    w += bonds_virial(s);
    if (s.coulomb == 1) {
        ewald_prepare(s, s.ewald);
        M3 real = ewald_real_virial(s, s.ewald);
        M3 k = ewald_k_virial(s, s.ewald);
        real += k;
        w += real;
    } else if (s.coulomb == 2) {
        w += wolf_virial(s, s.wolf);
    }
    return w;
}

// --- MolecularVirial (compute.rs:278-364) ------------------------------------------
// molecules.rs:256-263: centre of mass, masses and positions summed in atom order
V3 center_of_mass(const System& s, size_t m) {
    double total = 0.0;
    V3 com;
    for (size_t i = s.mol_start[m]; i < s.mol_start[m + 1]; i++) {
        total += s.mass[i];
        com += s.mass[i] * s.pos[i];
    }
    return com / total;
}

// w_ab * (r_ab . r_ij) / |r_ab|^2 for one atom pair of two molecules
inline M3 molecular_weight(const M3& w_ab, V3 r_ab, V3 r_ij) {
    double d = dot(r_ab, r_ij);
    double n2 = norm2(r_ab);
    M3 r;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) r.m[a][b] = w_ab.m[a][b] * d / n2;
    return r;
}

// Loop of compute.rs:285-310 (pairs), ewald.rs:533-570 (real space) and
// wolf.rs:285-324 over molecule pairs i < j and their atoms; `pair` returns
// the virial w_ab of one atom pair.
template <class F>
M3 molecule_pairs_virial(const System& s, F&& pair) {
    size_t nmol = s.mol_start.size() - 1;
    std::vector<V3> com(nmol);
    for (size_t m = 0; m < nmol; m++) com[m] = center_of_mass(s, m);
    int nt = std::max(1, s.nthreads);
    std::vector<M3> partial(nt);
    parallel_for_i(nt, nmol, [&](int t, size_t i0, size_t n_, size_t step) {
        M3 acc;
        for (size_t i = i0; i < n_; i += step) {
            M3 local;
            for (size_t j = i + 1; j < nmol; j++) {
                V3 r_ij = com[i] - com[j];
                s.vector_image(r_ij);
                for (size_t a = s.mol_start[i]; a < s.mol_start[i + 1]; a++)
                    for (size_t b = s.mol_start[j]; b < s.mol_start[j + 1]; b++) {
                        V3 r_ab = s.nearest_image(a, b);
                        M3 w;
                        if (pair(a, b, r_ab, w)) local += molecular_weight(w, r_ab, r_ij);
                    }
            }
            acc += local;
        }
        partial[t] = acc;
    });
    M3 w;
    for (auto& p : partial) w += p;
    return w;
}

M3 molecular_virial(System& s) {
    // pair potentials (compute.rs:285-310)
    M3 w = molecule_pairs_virial(s, [&](size_t a, size_t b, V3 r_ab, M3& out) {
        const PairInteraction* p = s.pair_potential(a, b);
        if (!p) return false;
        RInfo info = information(p->restriction, p->scale14, s.bond_path(a, b));
        if (info.excluded) return false;
        out = scale(info.scaling, p->virial(r_ab));
        return true;
    });
    // tails (compute.rs:313-322), as the atomic virial's
    double volume = s.volume();
    auto comp = composition(s);
    for (auto& [ki, ni] : comp)
        for (auto& [kj, nj] : comp) {
            double two_pi_density = 2.0 * PI * ni * nj / volume;
            auto it = s.pairs.find({std::min(ki, kj), std::max(ki, kj)});
            if (it != s.pairs.end()) {
                double diag = it->second.tail_virial_scalar() * (1.0 / 3.0);
                for (int k = 0; k < 3; k++) w.m[k][k] += two_pi_density * diag;
            }
        }
    // bonds (compute.rs:325-345) only warn; angles and dihedrals do not enter
    if (s.coulomb == 1) {
        Ewald& ew = s.ewald;
        ewald_prepare(s, ew);
        // ewald.rs:533-570: every charged pair, restriction info for the real-space force
        M3 real = molecule_pairs_virial(s, [&](size_t a, size_t b, V3 r_ab, M3& out) {
            double qa = s.charge[a], qb = s.charge[b];
            if (qa == 0.0 || qb == 0.0) return false;
            RInfo info = information(ew.restriction, ew.scale14, s.bond_path(a, b));
            V3 force = ewald_real_force_pair(ew, info, qa * qb, norm(r_ab)) * r_ab;
            out = tensorial(force, r_ab);
            return true;
        });
        // ewald.rs:736-753: atomic k-space virial minus sum_i F_i (x) (r_i - com)
        M3 k = ewald_k_virial(s, ew);
        std::vector<V3> fk(s.n);
        ewald_k_forces(s, ew, fk);
        M3 corr;
        for (size_t m = 0; m + 1 < s.mol_start.size(); m++) {
            V3 com = center_of_mass(s, m);
            for (size_t i = s.mol_start[m]; i < s.mol_start[m + 1]; i++) corr += tensorial(fk[i], s.pos[i] - com);
        }
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) k.m[a][b] = k.m[a][b] - corr.m[a][b];
        real += k;
        w += real;
    } else if (s.coulomb == 2) {
        const Wolf& wf = s.wolf;
        w += molecule_pairs_virial(s, [&](size_t a, size_t b, V3 r_ab, M3& out) {
            double qa = s.charge[a], qb = s.charge[b];
            if (qa == 0.0 || qb == 0.0) return false;
            RInfo info = information(wf.restriction, wf.scale14, s.bond_path(a, b));
            if (info.excluded) return false;
            V3 force = (info.scaling * wolf_force_pair(wf, qa * qb, norm(r_ab))) * r_ab;
            out = tensorial(force, r_ab);
            return true;
        });
    }
    return w;
}

double kinetic_energy(const System& s) {
    double e = 0.0;
    for (size_t i = 0; i < s.n; i++) e += 0.5 * s.mass[i] * norm2(s.vel[i]);
    return e;
}

double temperature(const System& s) {
    double dof = double(3 * s.n);  // system.rs:249-255 (particles)
    return 2.0 * kinetic_energy(s) / (dof * K_BOLTZMANN);
}

// bonding.rs:78-155: angles and dihedrals from bonds, and bond-distance flags
int rebuild_topology(System& s) {
    s.angles.clear();
    s.dihedrals.clear();
    s.bond_flags.clear();
    // molecules from connected components; must be contiguous ranges
    std::vector<size_t> parent(s.n);
    for (size_t i = 0; i < s.n; i++) parent[i] = i;
    std::function<size_t(size_t)> find = [&](size_t x) {
        while (parent[x] != x) x = parent[x] = parent[parent[x]];
        return x;
    };
    for (auto& [i, j] : s.bonds) {
        size_t a = find(i), b = find(j);
        if (a != b) parent[std::max(a, b)] = std::min(a, b);
    }
    s.molid.assign(s.n, 0);
    s.mol_start.clear();
    uint32_t m = 0;
    for (size_t i = 0; i < s.n; i++) {
        size_t root = find(i);
        if (root == i) {
            s.mol_start.push_back(i);
            s.molid[i] = m++;
        } else {
            if (s.molid[root] != m - 1 || (i > 0 && find(i - 1) != root)) {
                g_error = "molecules must be contiguous ranges of atoms";
                return 1;
            }
            s.molid[i] = s.molid[root];
        }
    }
    s.mol_start.push_back(s.n);

    std::set<std::tuple<size_t, size_t, size_t>> angles;
    std::set<std::tuple<size_t, size_t, size_t, size_t>> dihedrals;
    auto make_angle = [](size_t a, size_t b, size_t c) {
        return std::make_tuple(std::min(a, c), b, std::max(a, c));
    };
    auto make_dihedral = [](size_t a, size_t b, size_t c, size_t d) {
        if (std::max(a, b) < std::max(c, d)) return std::make_tuple(a, b, c, d);
        return std::make_tuple(d, c, b, a);
    };
    std::vector<std::pair<size_t, size_t>> bonds;
    for (auto& [i, j] : s.bonds) bonds.push_back({std::min(i, j), std::max(i, j)});
    for (auto& b1 : bonds) {
        for (auto& b2 : bonds) {
            if (b1 == b2) continue;
            std::tuple<size_t, size_t, size_t> angle;
            if (b1.first == b2.second)
                angle = make_angle(b2.first, b2.second, b1.second);
            else if (b1.second == b2.first)
                angle = make_angle(b1.first, b1.second, b2.second);
            else if (b1.second == b2.second)
                angle = make_angle(b1.first, b1.second, b2.first);
            else if (b1.first == b2.first)
                angle = make_angle(b1.second, b1.first, b2.second);
            else
                continue;
            angles.insert(angle);
            auto [ai, aj, ak] = angle;
            for (auto& b3 : bonds) {
                if (b2 == b3) continue;
                if (ak == b3.first && aj != b3.second)
                    dihedrals.insert(make_dihedral(ai, aj, ak, b3.second));
                else if (ak == b3.second && aj != b3.first)
                    dihedrals.insert(make_dihedral(ai, aj, ak, b3.first));
                else if (ai == b3.second && aj != b3.first)
                    dihedrals.insert(make_dihedral(b3.first, ai, aj, ak));
                else if (ai == b3.first && aj != b3.second)
                    dihedrals.insert(make_dihedral(b3.second, ai, aj, ak));
            }
        }
    }
    for (auto& a : angles) s.angles.push_back(a);
    for (auto& d : dihedrals) s.dihedrals.push_back(d);
    for (auto& [i, j] : bonds) {
        s.bond_flags[{i, j}] |= 1;
        s.bond_flags[{j, i}] |= 1;
    }
    for (auto& [i, j, k] : s.angles) {
        s.bond_flags[{i, k}] |= 2;
        s.bond_flags[{k, i}] |= 2;
    }
    for (auto& [i, j, k, m2] : s.dihedrals) {
        s.bond_flags[{i, m2}] |= 4;
        s.bond_flags[{m2, i}] |= 4;
    }
    return 0;
}

// --- MD (integrators.rs) ----------------------------------------------------------
void md_step_vv(System& s, double dt) {
    for (size_t i = 0; i < s.n; i++) {
        s.vel[i] += (0.5 * dt) * s.acc[i];
        s.pos[i] += s.vel[i] * dt;
    }
    std::vector<V3> f;
    all_forces(s, f);
    for (size_t i = 0; i < s.n; i++) s.acc[i] = f[i] / s.mass[i];
    for (size_t i = 0; i < s.n; i++) s.vel[i] += (0.5 * dt) * s.acc[i];
}

void md_step_leapfrog(System& s, double dt) {
    double dt2 = dt * dt;
    for (size_t i = 0; i < s.n; i++) s.pos[i] += s.vel[i] * dt + (0.5 * s.acc[i]) * dt2;
    std::vector<V3> f;
    all_forces(s, f);
    for (size_t i = 0; i < s.n; i++) {
        V3 na = f[i] / s.mass[i];
        s.vel[i] += (0.5 * (s.acc[i] + na)) * dt;
        s.acc[i] = na;
    }
}

// thermostats.rs:113-121
void berendsen(System& s, double T0, double tau) {
    double T = temperature(s);
    double factor = std::sqrt(1.0 + (T0 / T - 1.0) / tau);
    for (size_t i = 0; i < s.n; i++) s.vel[i] = s.vel[i] * factor;
}

// --- MC energy cache (cache.rs:75-220) ------------------------------------------
void cache_init(System& s) {
    size_t n = s.n;
    s.pair_cache.assign(n * n, 0.0);
    s.c_pairs = 0.0;
    for (size_t i = 0; i < n; i++)
        for (size_t j = i + 1; j < n; j++) {
            double r = norm(s.nearest_image(i, j));
            double e = pair_energy_of(s, s.bond_path(i, j), r, i, j);
            s.pair_cache[i * n + j] = e;
            s.pair_cache[j * n + i] = e;
            s.c_pairs += e;
        }
    s.c_tail = pairs_tail(s);
    s.c_bonds = bonds_energy(s);
    s.c_angles = angles_energy(s) + dihedrals_energy(s);
    s.c_coulomb = coulomb_energy(s);
    s.cache_ready = true;
    s.staged = false;
}

// ewald.rs:756-839 (delta rho of the moved molecule) and 572-617 (real part)
double ewald_move_cost(System& s, size_t molid, const std::vector<V3>& newpos) {
    Ewald& ew = s.ewald;
    ewald_prepare(s, ew);
    size_t m0 = s.mol_start[molid], m1 = s.mol_start[molid + 1];
    // real space
    double old_e = 0.0, new_e = 0.0;
    for (size_t a = m0; a < m1; a++) {
        double qi = s.charge[a];
        if (qi == 0.0) continue;
        for (size_t j = 0; j < s.n; j++) {
            if (j >= m0 && j < m1) continue;
            double qj = s.charge[j];
            if (qj == 0.0) continue;
            double old_r = s.distance(a, j);
            double new_r = s.cell_distance(newpos[a - m0], s.pos[j]);
            RInfo info = information(ew.restriction, ew.scale14, s.bond_path(a, j));
            old_e += ewald_real_energy_pair(ew, info, qi * qj, old_r);
            new_e += ewald_real_energy_pair(ew, info, qi * qj, new_r);
        }
    }
    double real = new_e - old_e;
    // k space: rho must be current (computed by cache_init / previous updates)
    if (ew.rho.size() != ew.factors.size()) {
        EikrTable t;
        eik_dot_r(s, ew, s.pos, t);
        compute_rho(s, ew, t, ew.rho);
    }
    double old_k = 0.0;
    for (size_t k = 0; k < ew.factors.size(); k++) old_k += ew.factors[k].energy * cnorm2(ew.rho[k]);
    old_k /= FOUR_PI_EPSILON_0;
    std::vector<V3> oldpos(s.pos.begin() + m0, s.pos.begin() + m1);
    EikrTable told, tnew;
    eik_dot_r(s, ew, oldpos, told);
    eik_dot_r(s, ew, newpos, tnew);
    ew.staged_delta.clear();
    for (auto& f : ew.factors) {
        Cx partial;
        for (size_t a = 0; a < m1 - m0; a++) {
            Cx op = cmul(cmul(told.at(f.ikx, 0, a), told.at(f.iky, 1, a)), told.at(f.ikz, 2, a));
            Cx np = cmul(cmul(tnew.at(f.ikx, 0, a), tnew.at(f.iky, 1, a)), tnew.at(f.ikz, 2, a));
            double q = s.charge[m0 + a];
            partial.re += q * (np.re - op.re);
            partial.im += q * (np.im - op.im);
        }
        ew.staged_delta.push_back(partial);
    }
    double new_k = 0.0;
    for (size_t k = 0; k < ew.factors.size(); k++) {
        Cx r{ew.rho[k].re + ew.staged_delta[k].re, ew.rho[k].im + ew.staged_delta[k].im};
        new_k += ew.factors[k].energy * cnorm2(r);
    }
    new_k /= FOUR_PI_EPSILON_0;
    ew.has_staged = true;
    return real + (new_k - old_k);
}

// wolf.rs:121-165
double wolf_move_cost(System& s, size_t molid, const std::vector<V3>& newpos) {
    const Wolf& w = s.wolf;
    size_t m0 = s.mol_start[molid], m1 = s.mol_start[molid + 1];
    double old_e = 0.0, new_e = 0.0;
    for (size_t a = m0; a < m1; a++) {
        double qi = s.charge[a];
        if (qi == 0.0) continue;
        for (size_t j = 0; j < s.n; j++) {
            if (j >= m0 && j < m1) continue;
            double qj = s.charge[j];
            if (qj == 0.0) continue;
            RInfo info = information(w.restriction, w.scale14, s.bond_path(a, j));
            if (info.excluded) continue;
            double old_r = s.distance(a, j);
            double new_r = s.cell_distance(newpos[a - m0], s.pos[j]);
            old_e += info.scaling * wolf_energy_pair(w, qi * qj, old_r);
            new_e += info.scaling * wolf_energy_pair(w, qi * qj, new_r);
        }
    }
    return new_e - old_e;
}

double move_molecule_cost(System& s, size_t molid, const std::vector<V3>& newpos) {
    size_t n = s.n;
    size_t m0 = s.mol_start[molid], m1 = s.mol_start[molid + 1];
    s.staged_new_pairs.assign((m1 - m0) * n, 0.0);
    double pairs_delta = 0.0;
    for (size_t a = m0; a < m1; a++) {
        for (size_t j = 0; j < n; j++) {
            if (j >= m0 && j < m1) continue;
            double r = s.cell_distance(s.pos[j], newpos[a - m0]);
            double e = pair_energy_of(s, s.bond_path(a, j), r, a, j);
            pairs_delta += e;
            s.staged_new_pairs[(a - m0) * n + j] += e;
            pairs_delta -= s.pair_cache[a * n + j];
        }
    }
    double coulomb_delta = 0.0;
    if (s.coulomb == 1) coulomb_delta = ewald_move_cost(s, molid, newpos);
    if (s.coulomb == 2) coulomb_delta = wolf_move_cost(s, molid, newpos);
    s.staged = true;
    s.staged_mol = molid;
    s.staged_pos = newpos;
    s.staged_pairs_delta = pairs_delta;
    s.staged_coulomb_delta = coulomb_delta;
    return pairs_delta + coulomb_delta;
}

// cache.rs:190-218 plus translate.rs:115-127 (positions, then COM wrap)
void move_commit(System& s, bool wrap_com) {
    size_t n = s.n;
    size_t m0 = s.mol_start[s.staged_mol], m1 = s.mol_start[s.staged_mol + 1];
    s.c_pairs += s.staged_pairs_delta;
    s.c_coulomb += s.staged_coulomb_delta;
    for (size_t a = m0; a < m1; a++) {
        for (size_t j = 0; j < n; j++) {
            if (j >= m0 && j < m1) continue;
            double e = s.staged_new_pairs[(a - m0) * n + j];
            s.pair_cache[a * n + j] = e;
            s.pair_cache[j * n + a] = e;
        }
    }
    if (s.coulomb == 1 && s.ewald.has_staged) {
        for (size_t k = 0; k < s.ewald.rho.size(); k++) {
            s.ewald.rho[k].re += s.ewald.staged_delta[k].re;
            s.ewald.rho[k].im += s.ewald.staged_delta[k].im;
        }
        s.ewald.has_staged = false;
    }
    for (size_t a = m0; a < m1; a++) s.pos[a] = s.staged_pos[a - m0];
    if (wrap_com && !s.infinite()) {
        // molecules.rs:256-296
        double total = 0.0;
        V3 com;
        for (size_t a = m0; a < m1; a++) {
            total += s.mass[a];
            com += s.mass[a] * s.pos[a];
        }
        com = com / total;
        V3 w = com;
        w.x -= std::floor(w.x / s.L[0]) * s.L[0];
        w.y -= std::floor(w.y / s.L[1]) * s.L[1];
        w.z -= std::floor(w.z / s.L[2]) * s.L[2];
        V3 delta = w - com;
        for (size_t a = m0; a < m1; a++) s.pos[a] += delta;
    }
    s.staged = false;
}

// --- Philox4x32-10 counter-based RNG (Salmon et al., SC'11) ----------------------
inline void philox_round(uint32_t ctr[4], uint32_t key[2]) {
    const uint64_t M0 = 0xD2511F53ull, M1 = 0xCD9E8D57ull;
    uint64_t p0 = M0 * ctr[0], p1 = M1 * ctr[2];
    uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
    uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    uint32_t n0 = hi1 ^ ctr[1] ^ key[0];
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ ctr[3] ^ key[1];
    uint32_t n3 = lo0;
    ctr[0] = n0;
    ctr[1] = n1;
    ctr[2] = n2;
    ctr[3] = n3;
}
void philox4x32_10(const uint32_t in[4], const uint32_t k[2], uint32_t out[4]) {
    uint32_t ctr[4] = {in[0], in[1], in[2], in[3]};
    uint32_t key[2] = {k[0], k[1]};
    for (int r = 0; r < 10; r++) {
        philox_round(ctr, key);
        if (r < 9) {
            key[0] += 0x9E3779B9u;
            key[1] += 0xBB67AE85u;
        }
    }
    for (int i = 0; i < 4; i++) out[i] = ctr[i];
}
// 53-bit uniform double in [0, 1) from two 32-bit words
inline double u01(uint32_t a, uint32_t b) {
    uint64_t x = (uint64_t(a) << 21) ^ uint64_t(b >> 11);
    x &= ((1ull << 53) - 1);
    return double(x) * (1.0 / 9007199254740992.0);
}

}  // namespace

// ============================================================================
// C ABI of the oracle (loaded with ctypes by the tests and bench.py)
// ============================================================================
extern "C" {

struct lo_system;

const char* lo_last_error() { return g_error.c_str(); }

lo_system* lo_create(size_t n) {
    System* s = new System();
    s->n = n;
    s->pos.assign(n, V3{});
    s->vel.assign(n, V3{});
    s->acc.assign(n, V3{});
    s->mass.assign(n, 1.0);
    s->charge.assign(n, 0.0);
    s->kind.assign(n, 0);
    s->molid.resize(n);
    for (size_t i = 0; i < n; i++) s->molid[i] = uint32_t(i);
    s->mol_start.resize(n + 1);
    for (size_t i = 0; i <= n; i++) s->mol_start[i] = i;
    s->nkinds = 1;
    return reinterpret_cast<lo_system*>(s);
}

void lo_destroy(lo_system* h) { delete reinterpret_cast<System*>(h); }

static System& S(lo_system* h) { return *reinterpret_cast<System*>(h); }

int lo_set_threads(lo_system* h, int n) {
    S(h).nthreads = n < 1 ? 1 : n;
    return 0;
}

int lo_set_cell(lo_system* h, double a, double b, double c) {
    System& s = S(h);
    s.L[0] = a;
    s.L[1] = b;
    s.L[2] = c;
    s.ewald.prev_a = -1;
    return 0;
}

int lo_set_positions(lo_system* h, const double* xyz) {
    System& s = S(h);
    for (size_t i = 0; i < s.n; i++) s.pos[i] = V3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    s.cache_ready = false;
    s.ewald.rho.clear();
    return 0;
}
int lo_get_positions(lo_system* h, double* xyz) {
    System& s = S(h);
    for (size_t i = 0; i < s.n; i++) {
        xyz[3 * i] = s.pos[i].x;
        xyz[3 * i + 1] = s.pos[i].y;
        xyz[3 * i + 2] = s.pos[i].z;
    }
    return 0;
}
int lo_set_velocities(lo_system* h, const double* v) {
    System& s = S(h);
    for (size_t i = 0; i < s.n; i++) s.vel[i] = V3{v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    return 0;
}
int lo_get_velocities(lo_system* h, double* v) {
    System& s = S(h);
    for (size_t i = 0; i < s.n; i++) {
        v[3 * i] = s.vel[i].x;
        v[3 * i + 1] = s.vel[i].y;
        v[3 * i + 2] = s.vel[i].z;
    }
    return 0;
}
int lo_set_masses(lo_system* h, const double* m) {
    System& s = S(h);
    for (size_t i = 0; i < s.n; i++) s.mass[i] = m[i];
    return 0;
}
int lo_set_charges(lo_system* h, const double* q) {
    System& s = S(h);
    for (size_t i = 0; i < s.n; i++) s.charge[i] = q[i];
    s.ewald.rho.clear();
    return 0;
}
int lo_set_kinds(lo_system* h, const uint32_t* k, uint32_t nkinds) {
    System& s = S(h);
    for (size_t i = 0; i < s.n; i++) {
        if (k[i] >= nkinds) {
            g_error = "kind index out of range";
            return 1;
        }
        s.kind[i] = k[i];
    }
    s.nkinds = nkinds;
    return 0;
}
int lo_set_bonds(lo_system* h, size_t nbonds, const uint32_t* ij) {
    System& s = S(h);
    s.bonds.clear();
    for (size_t b = 0; b < nbonds; b++) {
        size_t i = ij[2 * b], j = ij[2 * b + 1];
        if (i >= s.n || j >= s.n || i == j) {
            g_error = "invalid bond";
            return 1;
        }
        s.bonds.push_back({std::min(i, j), std::max(i, j)});
    }
    std::sort(s.bonds.begin(), s.bonds.end());
    s.bonds.erase(std::unique(s.bonds.begin(), s.bonds.end()), s.bonds.end());
    return rebuild_topology(s);
}

int lo_set_pair(lo_system* h, uint32_t ki, uint32_t kj, int pot, const double* params, double cutoff,
                int shifted, int tail, int restriction, double scale14) {
    System& s = S(h);
    PairInteraction p;
    if (!set_potential(p.pot, pot, params)) {
        g_error = "invalid pair potential parameters";
        return 1;
    }
    p.cutoff = cutoff;
    p.shifted = shifted != 0;
    p.shift = p.shifted ? p.pot.energy(cutoff) : 0.0;  // pairs.rs:86-95
    p.tail = tail != 0;
    p.restriction = restriction;
    p.scale14 = scale14;
    s.pairs[{std::min(ki, kj), std::max(ki, kj)}] = p;
    return 0;
}
// PairInteraction over TableComputation::new(potential, size, max)
// (computations.rs:96-117); the shift is the table's energy at the cutoff.
int lo_set_pair_table(lo_system* h, uint32_t ki, uint32_t kj, int pot, const double* params, uint64_t size,
                      double max, double cutoff, int shifted, int tail, int restriction, double scale14) {
    System& s = S(h);
    PairInteraction p;
    if (!set_potential(p.pot, pot, params) || size < 1 || !(max > 0.0)) {
        g_error = "invalid tabulated pair potential";
        return 1;
    }
    p.pot.make_table(size_t(size), max);
    p.cutoff = cutoff;
    p.shifted = shifted != 0;
    p.shift = p.shifted ? p.pot.energy(cutoff) : 0.0;
    p.tail = tail != 0;
    p.restriction = restriction;
    p.scale14 = scale14;
    s.pairs[{std::min(ki, kj), std::max(ki, kj)}] = p;
    return 0;
}
int lo_set_bond_potential(lo_system* h, uint32_t ki, uint32_t kj, int pot, const double* params) {
    Potential p;
    if (!set_potential(p, pot, params)) return 1;
    S(h).bond_pots[{std::min(ki, kj), std::max(ki, kj)}] = p;
    return 0;
}
int lo_set_angle_potential(lo_system* h, uint32_t ki, uint32_t kj, uint32_t kk, int pot, const double* params) {
    Potential p;
    if (!set_potential(p, pot, params)) return 1;
    S(h).angle_pots[{std::min(ki, kk), kj, std::max(ki, kk)}] = p;
    return 0;
}
int lo_set_dihedral_potential(lo_system* h, uint32_t ki, uint32_t kj, uint32_t kk, uint32_t km, int pot,
                              const double* params) {
    Potential p;
    if (!set_potential(p, pot, params)) return 1;
    S(h).dihedral_pots[System::dihedral_key(ki, kj, kk, km)] = p;
    return 0;
}
int lo_set_ewald(lo_system* h, double cutoff, int kmax, double alpha, int restriction, double scale14) {
    System& s = S(h);
    if (std::isnan(alpha)) alpha = PI / cutoff;  // ewald.rs:279
    if (cutoff < 0.0 || alpha < 0.0 || kmax <= 0) {
        g_error = "invalid Ewald parameters";
        return 1;
    }
    s.ewald = Ewald{};
    s.ewald.alpha = alpha;
    s.ewald.rc = cutoff;
    s.ewald.kmax = kmax;
    s.ewald.restriction = restriction;
    s.ewald.scale14 = scale14;
    s.coulomb = 1;
    return 0;
}
int lo_set_wolf(lo_system* h, double cutoff, int restriction, double scale14) {
    System& s = S(h);
    if (!(cutoff > 0.0)) {
        g_error = "Got a negative cutoff in Wolf summation";
        return 1;
    }
    s.wolf = Wolf{};
    wolf_init(s.wolf, cutoff);
    s.wolf.restriction = restriction;
    s.wolf.scale14 = scale14;
    s.coulomb = 2;
    return 0;
}

int lo_forces(lo_system* h, double* out) {
    System& s = S(h);
    std::vector<V3> f;
    all_forces(s, f);
    for (size_t i = 0; i < s.n; i++) {
        out[3 * i] = f[i].x;
        out[3 * i + 1] = f[i].y;
        out[3 * i + 2] = f[i].z;
    }
    return 0;
}

// terms: pairs, pairs_tail, bonds, angles, dihedrals, coulomb, global, total
int lo_energy(lo_system* h, double* terms) {
    System& s = S(h);
    double pairs = pairs_energy(s);
    double tail = pairs_tail(s);
    double bonds = bonds_energy(s);
    double angles = angles_energy(s);
    double dihedrals = dihedrals_energy(s);
    double coul = coulomb_energy(s);
    // compute.rs:114-127 accumulation order
    double total = pairs;
    total += tail;
    total += bonds;
    total += angles;
    total += dihedrals;
    total += coul;
    total += 0.0;
    terms[0] = pairs;
    terms[1] = tail;
    terms[2] = bonds;
    terms[3] = angles;
    terms[4] = dihedrals;
    terms[5] = coul;
    terms[6] = 0.0;
    terms[7] = total;
    return std::isfinite(total) ? 0 : 1;
}

// Ewald components: real, self, kspace energies; real and kspace virials (9 each)
int lo_ewald_components(lo_system* h, double* energies3, double* real_virial9, double* k_virial9) {
    System& s = S(h);
    if (s.coulomb != 1) {
        g_error = "no Ewald potential";
        return 1;
    }
    ewald_prepare(s, s.ewald);
    energies3[0] = ewald_real_energy(s, s.ewald);
    energies3[1] = ewald_self_energy(s, s.ewald);
    energies3[2] = ewald_k_energy(s, s.ewald);
    M3 wr = ewald_real_virial(s, s.ewald);
    M3 wk = ewald_k_virial(s, s.ewald);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            real_virial9[3 * a + b] = wr.m[a][b];
            k_virial9[3 * a + b] = wk.m[a][b];
        }
    return 0;
}

// number of k-vectors in the Ewald half-space set
int lo_ewald_kvectors(lo_system* h, int* out_count) {
    System& s = S(h);
    if (s.coulomb != 1) return 1;
    ewald_prepare(s, s.ewald);
    *out_count = int(s.ewald.factors.size());
    return 0;
}

// MolecularVirial (compute.rs:278-364)
int lo_molecular_virial(lo_system* h, double* w9) {
    System& s = S(h);
    if (s.infinite()) {
        g_error = "Can not compute virial for infinite cell";
        return 1;
    }
    M3 w = molecular_virial(s);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) w9[3 * a + b] = w.m[a][b];
    return 0;
}
int lo_virial(lo_system* h, double* w9) {
    System& s = S(h);
    M3 w = atomic_virial(s);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) w9[3 * a + b] = w.m[a][b];
    return 0;
}

double lo_kinetic_energy(lo_system* h) { return kinetic_energy(S(h)); }
double lo_temperature(lo_system* h) { return temperature(S(h)); }

// Count of pairs the cutoff loop keeps: pairs (r < rc_pair, not excluded) and
// charged pairs (r <= rc_coulomb).  Used for the algorithmic work per step.
int lo_pair_counts(lo_system* h, uint64_t* n_pair, uint64_t* n_coul) {
    System& s = S(h);
    uint64_t np = 0, nc = 0;
    double rcc = s.coulomb == 1 ? s.ewald.rc : (s.coulomb == 2 ? s.wolf.cutoff : -1.0);
    for (size_t i = 0; i < s.n; i++)
        for (size_t j = i + 1; j < s.n; j++) {
            double r = norm(s.nearest_image(i, j));
            const PairInteraction* p = s.pair_potential(i, j);
            if (p && p->pot.kind != POT_NULL) {
                RInfo info = information(p->restriction, p->scale14, s.bond_path(i, j));
                if (!info.excluded && r < p->cutoff) np++;
            }
            if (rcc > 0 && s.charge[i] != 0.0 && s.charge[j] != 0.0 && r <= rcc) nc++;
        }
    *n_pair = np;
    *n_coul = nc;
    return 0;
}

// The exact candidate set of the brute-force cutoff loop: every (i<j) with
// r < rc (strict), written as pairs; returns the count (buffer may be null).
uint64_t lo_pairs_within(lo_system* h, double rc, uint32_t* out, uint64_t capacity) {
    System& s = S(h);
    uint64_t c = 0;
    for (size_t i = 0; i < s.n; i++)
        for (size_t j = i + 1; j < s.n; j++) {
            double r = norm(s.nearest_image(i, j));
            if (r < rc) {
                if (out && c < capacity) {
                    out[2 * c] = uint32_t(i);
                    out[2 * c + 1] = uint32_t(j);
                }
                c++;
            }
        }
    return c;
}

// MD: integrator 0 = Velocity-Verlet, 1 = LeapFrog.  berendsen_T >= 0 applies
// BerendsenThermostat(berendsen_T, tau) after each step (T = 0 is allowed, as
// thermostats.rs:103 asserts temperature >= 0); a negative or NaN value
// disables it.  The integrator starts from zero accelerations on the first
// call after lo_md_reset (integrators.rs:40-42, :146-148).
int lo_md_reset(lo_system* h) {
    System& s = S(h);
    s.acc.assign(s.n, V3{});
    return 0;
}
int lo_md_run(lo_system* h, int integrator, double dt, uint64_t nsteps, double berendsen_T, double tau) {
    System& s = S(h);
    if (s.acc.size() != s.n) s.acc.assign(s.n, V3{});
    for (uint64_t st = 0; st < nsteps; st++) {
        if (integrator == 0)
            md_step_vv(s, dt);
        else
            md_step_leapfrog(s, dt);
        if (berendsen_T >= 0.0) berendsen(s, berendsen_T, tau);
    }
    return 0;
}

// MC (cache.rs): init the cache, cost of moving molecule `molid` to `newpos`,
// then commit (update + COM wrap as translate.rs:115-127) or discard.
int lo_cache_init(lo_system* h) {
    cache_init(S(h));
    return 0;
}
double lo_cache_energy(lo_system* h) {
    System& s = S(h);
    return s.c_pairs + s.c_tail + s.c_bonds + s.c_angles + s.c_coulomb;
}
int lo_move_cost(lo_system* h, uint32_t molid, const double* newpos, double* cost) {
    System& s = S(h);
    if (!s.cache_ready) cache_init(s);
    if (molid + 1 >= s.mol_start.size()) {
        g_error = "molecule id out of range";
        return 1;
    }
    size_t m0 = s.mol_start[molid], m1 = s.mol_start[molid + 1];
    std::vector<V3> np(m1 - m0);
    for (size_t a = 0; a < m1 - m0; a++) np[a] = V3{newpos[3 * a], newpos[3 * a + 1], newpos[3 * a + 2]};
    *cost = move_molecule_cost(s, molid, np);
    return 0;
}
int lo_move_commit(lo_system* h, int wrap_com) {
    System& s = S(h);
    if (!s.staged) {
        g_error = "no staged move";
        return 1;
    }
    move_commit(s, wrap_com != 0);
    return 0;
}
int lo_move_discard(lo_system* h) {
    System& s = S(h);
    s.staged = false;
    s.ewald.has_staged = false;
    return 0;
}

// Single-particle translation Metropolis chain (monte_carlo.rs:289-330,
// translate.rs:85-127) with a Philox4x32-10 stream keyed by `seed`.
// Each attempt t uses counter (t, 0, 0, 0): words 0,1 -> molecule choice,
// words 2,3 and a second block (t, 1, 0, 0) -> displacement and acceptance.
int lo_mc_run(lo_system* h, uint64_t seed, uint64_t attempt0, uint64_t nattempts, double beta, double delta,
              uint64_t* accepted) {
    System& s = S(h);
    if (!s.cache_ready) cache_init(s);
    size_t nmol = s.mol_start.size() - 1;
    double d = delta / std::sqrt(3.0);  // translate.rs:43
    uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    uint64_t acc = 0;
    for (uint64_t t = attempt0; t < attempt0 + nattempts; t++) {
        uint32_t c0[4] = {uint32_t(t), uint32_t(t >> 32), 0, 0};
        uint32_t c1[4] = {uint32_t(t), uint32_t(t >> 32), 1, 0};
        uint32_t r0[4], r1[4];
        philox4x32_10(c0, key, r0);
        philox4x32_10(c1, key, r1);
        size_t mol = size_t(u01(r0[0], r0[1]) * double(nmol));
        if (mol >= nmol) mol = nmol - 1;
        double ux = u01(r0[2], r0[3]), uy = u01(r1[0], r1[1]), uz = u01(r1[2], r1[3]);
        uint32_t c2[4] = {uint32_t(t), uint32_t(t >> 32), 2, 0};
        uint32_t r2[4];
        philox4x32_10(c2, key, r2);
        double uacc = u01(r2[0], r2[1]);
        V3 disp{-d + 2.0 * d * ux, -d + 2.0 * d * uy, -d + 2.0 * d * uz};
        size_t m0 = s.mol_start[mol], m1 = s.mol_start[mol + 1];
        std::vector<V3> np(m1 - m0);
        for (size_t a = m0; a < m1; a++) np[a - m0] = s.pos[a] + disp;
        double cost = beta * move_molecule_cost(s, mol, np);
        bool ok = cost <= 0.0 || uacc < std::exp(-cost);
        if (ok) {
            move_commit(s, true);
            acc++;
        } else {
            s.staged = false;
            s.ewald.has_staged = false;
        }
    }
    *accepted = acc;
    return 0;
}

void lo_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox4x32_10(ctr, key, out); }

// Scalar functions for the unit golden values.
// Scalar values of one potential (energy/force at r, tails at rc), and of
// its TableComputation when size > 0; NaN for invalid parameters.
double lo_potential_value(int pot, const double* params, uint64_t size, double max, int what, double x) {
    Potential p;
    if (!set_potential(p, pot, params)) return std::nan("");
    if (size > 0) p.make_table(size_t(size), max);
    switch (what) {
        case 0: return p.energy(x);
        case 1: return p.force(x);
        case 2: return p.tail_energy(x);
        case 3: return p.tail_virial(x);
    }
    return std::nan("");
}
double lo_lj_energy(double sigma, double eps, double r) {
    Potential p;
    p.kind = POT_LJ;
    p.p0 = sigma;
    p.p1 = eps;
    return p.energy(r);
}
double lo_lj_force(double sigma, double eps, double r) {
    Potential p;
    p.kind = POT_LJ;
    p.p0 = sigma;
    p.p1 = eps;
    return p.force(r);
}
double lo_lj_tail_energy(double sigma, double eps, double rc) {
    Potential p;
    p.kind = POT_LJ;
    p.p0 = sigma;
    p.p1 = eps;
    return p.tail_energy(rc);
}
double lo_lj_tail_virial(double sigma, double eps, double rc) {
    Potential p;
    p.kind = POT_LJ;
    p.p0 = sigma;
    p.p1 = eps;
    return p.tail_virial(rc);
}
double lo_pair_interaction_energy(int pot, double p0, double p1, double rc, int shifted, double r) {
    PairInteraction pi;
    pi.pot.kind = pot;
    pi.pot.p0 = p0;
    pi.pot.p1 = p1;
    pi.cutoff = rc;
    pi.shifted = shifted != 0;
    pi.shift = pi.shifted ? pi.pot.energy(rc) : 0.0;
    return pi.energy(r);
}
double lo_pair_interaction_force(int pot, double p0, double p1, double rc, double r) {
    PairInteraction pi;
    pi.pot.kind = pot;
    pi.pot.p0 = p0;
    pi.pot.p1 = p1;
    pi.cutoff = rc;
    return pi.force(r);
}

}  // extern "C"
