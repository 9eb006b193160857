// Host-only check of the per-tile launch orders of rhybrid_amd/csrc/rh_geom.hpp:
// for many tile-grid shapes, the slab order (xcd_tile) and the brick order
// (xcd_brick_tile), over the 8 * xchunk workgroups of a per-tile launch,
// visit every tile exactly once; xcd_index does the same for any count.
// Prints "ok" and exits 0, or names the first failing shape and exits 1.
#include <cstdio>
#include <vector>

#include "../rhybrid_amd/csrc/rh_geom.hpp"

using namespace rh;

static bool covers(const Geom& g, int (*order)(const Geom&, int)) {
    std::vector<int> seen(g.ntiles, 0);
    for (int b = 0; b < N_XCD * g.xchunk; ++b) {
        const int t = order(g, b);
        if (t < 0) return false;
        if (t >= g.ntiles) continue;
        if (seen[t]++) return false;
    }
    for (int v : seen)
        if (v != 1) return false;
    return true;
}

int main() {
    for (int nx = 1; nx <= 13; ++nx)
        for (int ny = 1; ny <= 11; ++ny)
            for (int nz = 1; nz <= 9; ++nz) {
                Geom g{};
                g.ntx = nx; g.nty = ny; g.ntz = nz;
                g.ntiles = nx * ny * nz;
                xcd_bricks(g);
                if (g.xpx * g.xpy * g.xpz != N_XCD || !covers(g, xcd_tile) || !covers(g, xcd_brick_tile)) {
                    std::printf("tile grid %d x %d x %d: order does not cover every tile once\n", nx, ny, nz);
                    return 1;
                }
            }
    const int shapes[][3] = {{32, 24, 24}, {64, 48, 48}, {16, 12, 12}, {1, 1, 64}};
    for (const auto& s : shapes) {
        Geom g{};
        g.ntx = s[0]; g.nty = s[1]; g.ntz = s[2];
        g.ntiles = s[0] * s[1] * s[2];
        xcd_bricks(g);
        if (!covers(g, xcd_tile) || !covers(g, xcd_brick_tile)) {
            std::printf("tile grid %d x %d x %d: order does not cover every tile once\n", s[0], s[1], s[2]);
            return 1;
        }
    }
    for (int n = 1; n <= 5000; n += 7) {
        std::vector<int> seen(n, 0);
        const int grid = N_XCD * ((n + N_XCD - 1) / N_XCD);
        for (int b = 0; b < grid; ++b) {
            const int i = xcd_index(n, b);
            if (i < n && seen[i]++) { std::printf("xcd_index %d repeats\n", n); return 1; }
        }
        for (int v : seen)
            if (v != 1) { std::printf("xcd_index %d misses\n", n); return 1; }
    }
    std::printf("ok\n");
    return 0;
}
