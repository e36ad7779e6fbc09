// Host-side bounds check of the index arithmetic that the device kernels
// derive from coordinates: cell indices of the neighbour-list builds
// (cell_coord, also behind the direct build's cell tags), the Monte Carlo
// cell lists (mc_axis_cell) and the piece index of the erfcx tables.  The
// functions are the kernels' own (__host__ __device__ in the headers),
// evaluated on the CPU for finite, huge, infinite and NaN inputs: every
// index must lie in range, so that a simulation whose coordinates blow up
// keeps every kernel's accesses in bounds.  Prints "ok" and exits 0, or
// names the first failure and exits 1.
#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

#include "../../lumol_amd/csrc/lg_mc.hpp"

using namespace lg;

int main() {
    const double inf = std::numeric_limits<double>::infinity();
    const double nan = std::numeric_limits<double>::quiet_NaN();
    std::vector<double> xs = {0.0, -0.0, 1e-300, -1e-300, 0.5, 7.9, 8.0, 15.999999999999998, 16.0, 16.000000001,
                              -16.0, -1e-12, 1e6, -1e6, 1e15, -1e15, 1e300, -1e300, 1.7976931348623157e308,
                              -1.7976931348623157e308, inf, -inf, nan, -nan};
    const double Ls[] = {16.0, 44.8, 1000.0, 3.0};
    const int ncs[] = {1, 2, 3, 5, 37, 256, 1023};
    long checks = 0;
    for (double L : Ls)
        for (int nc : ncs)
            for (int periodic = 0; periodic < 2; periodic++)
                for (double x : xs) {
                    const double inv_cs = nc / L;
                    const int c = cell_coord(x, L, 1.0 / L, nc, inv_cs, periodic);
                    if (c < 0 || c >= nc) {
                        std::printf("cell_coord(%g, L=%g, nc=%d, periodic=%d) = %d\n", x, L, nc, periodic, c);
                        return 1;
                    }
                    const int m = mc_axis_cell(x, L, 1.0 / L, inv_cs, nc);
                    if (m < 0 || m >= nc) {
                        std::printf("mc_axis_cell(%g, L=%g, nc=%d) = %d\n", x, L, nc, m);
                        return 1;
                    }
                    checks += 2;
                }
    // cell tags: every axis field within the grid
    Grid g{};
    g.n[0] = 37;
    g.n[1] = 5;
    g.n[2] = 1023;
    Box b{};
    for (int d = 0; d < 3; d++) {
        b.L[d] = 44.8;
        b.inv[d] = 1.0 / 44.8;
        g.size[d] = 44.8 / g.n[d];
        g.inv_size[d] = g.n[d] / 44.8;
    }
    b.periodic = 1;
    CellTag t{nullptr, g, b};
    for (double x : xs)
        for (double y : xs) {
            const int tag = cell_tag(t, make_double4(x, y, nan, 0.0));
            const int cx = tag & 1023, cy = (tag >> 10) & 1023, cz = tag >> 20;
            if (cx >= g.n[0] || cy >= g.n[1] || cz >= g.n[2] || tag < 0) {
                std::printf("cell_tag(%g, %g, nan) = %d\n", x, y, tag);
                return 1;
            }
            checks++;
        }
    // erfcx piece index: the table has kErfcxPieces * (deg + 1) entries;
    // a table of that size whose every entry is 1 returns a finite value only
    // if the read stays inside (the check also reads it through the index)
    std::vector<double> t64(size_t(kErfcxPieces) * (kErfcxDeg64 + 1), 0.0);
    std::vector<float> t32(size_t(kErfcxPieces) * (kErfcxDeg32 + 1), 0.0f);
    for (double x : xs) {
        for (double inv_h : {1.0, 85.3, 1e300}) {
            const double u = x * inv_h;
            const int k64 = (int)fmin(fmax(u, 0.0), (double)(kErfcxPieces - 1));
            const float uf = (float)x * (float)inv_h;
            const int k32 = (int)fminf(fmaxf(uf, 0.f), (float)(kErfcxPieces - 1));
            if (k64 < 0 || k64 >= kErfcxPieces || k32 < 0 || k32 >= kErfcxPieces) {
                std::printf("erfcx piece index out of range for x=%g\n", x);
                return 1;
            }
            (void)erfcx_tab(t64.data(), x, inv_h);
            (void)erfcx_tab(t32.data(), (float)x, (float)inv_h);
            checks++;
        }
    }
    std::printf("ok %ld\n", checks);
    return 0;
}
