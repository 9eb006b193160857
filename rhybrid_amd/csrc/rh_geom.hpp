// rh_geom.hpp -- grid and tile geometry shared by host and device code.
//
// Cells are numbered as in the reference, n = i + nx*(j + ny*k) (ParGrid
// global ids with x fastest, src/particle_list_hybrid.h:89-91).  Particles are
// kept sorted by *subcell key* in tile-major order:
//   key = ((tile * tc) + local_cell) * 8 + octant
// where the grid is cut into tiles of tx*ty*tz cells, local_cell is the cell's
// index inside its tile (x fastest) and the octant says in which half of the
// cell the particle sits along each axis (bit0: x >= dx/2, bit1: y, bit2: z).
// A tile's particles are one contiguous range, each cell's particles are one
// contiguous range, and each subcell's particles share one 2x2x2 cloud-in-
// cell footprint, which is what the deposit exploits.
//
// The "extended region" of a tile is the tile grown by one cell on every side,
// (tx+2)*(ty+2)*(tz+2) cells with local coordinates -1..t.  A particle pushed
// out of a cell of the tile lands in the extended region (moves are at most
// one cell per step), and a deposit from the tile's subcells only touches it.
#pragma once

#include <cstdint>

#ifndef RH_HD
#if defined(__HIPCC__)
#define RH_HD __host__ __device__ __forceinline__
#else
#define RH_HD inline
#endif
#endif

namespace rh {

constexpr uint32_t DEAD_CELL = 0xFFFFFFFFu;  // particle marked for removal
constexpr uint32_t DEAD_PACKED = 0xFFFFFFFFu;
constexpr int RANK_BITS = 20;
constexpr uint32_t RANK_MASK = (1u << RANK_BITS) - 1u;
constexpr int MAX_EXTK = 4095;               // ext cells * 8 must fit 12 bits (4095 = dead)

struct Geom {
    int nx, ny, nz, nc;
    int per[3];
    double dx;
    double x0, y0, z0;
    int tx, ty, tz, tc;        // tile shape (powers of two), cells per tile
    int ltx, lty, ltz;         // log2 of the tile shape
    int ntx, nty, ntz, ntiles; // tiles per axis
    int ex, ey, ez, next;      // extended region shape and cell count
    int extk;                  // next * 8 subcell slots in the extended region
    uint32_t nkeys;            // ntiles * tc * 8
    int xpx, xpy, xpz, xchunk; // XCD bricks: parts per axis (product 8), most tiles in one brick
};

RH_HD void cell_ijk(const Geom& g, int n, int& i, int& j, int& k) {
    i = n % g.nx;
    j = (n / g.nx) % g.ny;
    k = n / (g.nx * g.ny);
}

RH_HD int cell_index(const Geom& g, int i, int j, int k) { return i + g.nx * (j + g.ny * k); }

// wrap a coordinate along one axis; -1 if it leaves a non-periodic axis
RH_HD int wrap_axis(int i, int n, int per) {
    if (i >= 0 && i < n) return i;
    if (!per) return -1;
    i %= n;
    return i < 0 ? i + n : i;
}

// neighbour at offset (a,b,c) of cell (i,j,k); -1 if it does not exist
RH_HD int nbr(const Geom& g, int i, int j, int k, int a, int b, int c) {
    const int ii = wrap_axis(i + a, g.nx, g.per[0]);
    const int jj = wrap_axis(j + b, g.ny, g.per[1]);
    const int kk = wrap_axis(k + c, g.nz, g.per[2]);
    if (ii < 0 || jj < 0 || kk < 0) return -1;
    return cell_index(g, ii, jj, kk);
}

// all 26 neighbours exist (pargrid::ALL_NEIGHBOURS_EXIST)
RH_HD bool all_nbrs(const Geom& g, int i, int j, int k) {
    if (!g.per[0] && (i < 1 || i > g.nx - 2)) return false;
    if (!g.per[1] && (j < 1 || j > g.ny - 2)) return false;
    if (!g.per[2] && (k < 1 || k > g.nz - 2)) return false;
    return true;
}

RH_HD bool has_pos(int i, int n, int per) { return per || i < n - 1; }
RH_HD bool has_neg(int i, int per) { return per || i > 0; }

// the cell is missing a +x, +y or +z neighbour (node and face kernels skip it)
RH_HD bool upper_missing(const Geom& g, int i, int j, int k) {
    return !has_pos(i, g.nx, g.per[0]) || !has_pos(j, g.ny, g.per[1]) || !has_pos(k, g.nz, g.per[2]);
}

RH_HD int tile_of(const Geom& g, int i, int j, int k) {
    return (i >> g.ltx) + g.ntx * ((j >> g.lty) + g.nty * (k >> g.ltz));
}

// local cell index inside a tile -> local coordinates
RH_HD void local_ijk(const Geom& g, int lc, int& li, int& lj, int& lk) {
    li = lc & (g.tx - 1);
    lj = (lc >> g.ltx) & (g.ty - 1);
    lk = lc >> (g.ltx + g.lty);
}

// XCD-aware tile orders.  The hardware deals the workgroups of a launch to
// the eight XCDs (each with its own L2) in turn, workgroup b to XCD b % 8.
// Per-tile launches use a 1D grid of 8 * xchunk workgroups (xchunk covers
// both orders); a result >= ntiles is no tile.
//   xcd_tile: XCD x takes the contiguous range [x * chunk, (x + 1) * chunk)
//   of the x-fastest tile order (slabs of whole tile layers), so that tiles
//   neighbouring in memory run on one XCD close in time.  The push and the
//   deposit, which read each tile's particle range, run in this order.
//   xcd_brick_tile: the tile grid cut into eight bricks (xpx x xpy x xpz
//   parts per axis), XCD x taking brick x, x-fastest.  The move writes the
//   particles that cross into a neighbour tile into that tile's subcell
//   ranges, in cache lines the neighbour's own workgroup also writes; within
//   a brick those two workgroups share an L2 and run closer in time than in
//   slabs, and bricks share fewer faces between XCDs (the move 519 -> 500 us
//   per launch on the 128x96x96 box; the deposit and the push were slower
//   in this order, 289 -> 307 us).
constexpr int N_XCD = 8;
RH_HD int xcd_tile(const Geom& g, int b) {
    const int chunk = (g.ntiles + N_XCD - 1) / N_XCD;
    const int i = b / N_XCD;
    return i < chunk ? (b % N_XCD) * chunk + i : g.ntiles;
}

RH_HD int xcd_brick_tile(const Geom& g, int b) {
    const int x = b % N_XCD, i = b / N_XCD;
    const int ix = x % g.xpx, iy = (x / g.xpx) % g.xpy, iz = x / (g.xpx * g.xpy);
    const int x0 = ix * g.ntx / g.xpx, x1 = (ix + 1) * g.ntx / g.xpx;
    const int y0 = iy * g.nty / g.xpy, y1 = (iy + 1) * g.nty / g.xpy;
    const int z0 = iz * g.ntz / g.xpz, z1 = (iz + 1) * g.ntz / g.xpz;
    const int bx = x1 - x0, by = y1 - y0, bz = z1 - z0;
    if (i >= bx * by * bz) return g.ntiles;
    return (x0 + i % bx) + g.ntx * ((y0 + (i / bx) % by) + g.nty * (z0 + i / (bx * by)));
}

// The brick split: eight parts, each doubling of an axis's parts going to
// the axis with the most tiles per part (z first on ties).
inline void xcd_bricks(Geom& g) {
    int p[3] = {1, 1, 1};
    const int n[3] = {g.ntx, g.nty, g.ntz};
    for (int s = 0; s < 3; ++s) {
        int best = 2;
        for (int d = 1; d >= 0; --d)
            if (n[d] * p[best] > n[best] * p[d]) best = d;  // n[d]/p[d] > n[best]/p[best]
        p[best] *= 2;
    }
    g.xpx = p[0]; g.xpy = p[1]; g.xpz = p[2];
    const int brick = ((g.ntx + p[0] - 1) / p[0]) * ((g.nty + p[1] - 1) / p[1]) * ((g.ntz + p[2] - 1) / p[2]);
    const int slab = (g.ntiles + N_XCD - 1) / N_XCD;
    g.xchunk = brick > slab ? brick : slab;
}

// The same dealing for any count of tiles (n), e.g. the field update's.
RH_HD int xcd_index(int n, int b) {
    const int chunk = (n + N_XCD - 1) / N_XCD;
    return (b % N_XCD) * chunk + b / N_XCD;
}

RH_HD void tile_coords(const Geom& g, int T, int& bx, int& by, int& bz) {
    bx = T % g.ntx;
    by = (T / g.ntx) % g.nty;
    bz = T / (g.ntx * g.nty);
}

RH_HD void tile_origin(const Geom& g, int T, int& ox, int& oy, int& oz) {
    ox = (T % g.ntx) * g.tx;
    oy = ((T / g.ntx) % g.nty) * g.ty;
    oz = (T / (g.ntx * g.nty)) * g.tz;
}

RH_HD uint32_t cell_key(const Geom& g, int i, int j, int k) {
    const int T = tile_of(g, i, j, k);
    const int lc = (i & (g.tx - 1)) + ((j & (g.ty - 1)) << g.ltx) + ((k & (g.tz - 1)) << (g.ltx + g.lty));
    return (uint32_t)(T * g.tc + lc);
}

// Candidate (tile index along the axis, extended coordinate 0..t+1) pairs
// whose extended region holds coordinate i of an axis of n cells cut into nt
// tiles of t = 2^lt cells.  Visited in a fixed order: tile offset -1,0,+1, then
// representative -n,0,+n (the same tile can hold a periodic cell twice when
// there are fewer than three tiles along the axis).  Returns the count (<=6).
RH_HD int axis_candidates(int i, int n, int lt, int nt, int per, int* tl, int* el) {
    const int t = 1 << lt;
    int m = 0;
    const int ti = i >> lt;
    int seen[3];
    int ns = 0;
    for (int d = -1; d <= 1; ++d) {
        int tt = ti + d;
        if (tt < 0 || tt >= nt) {
            if (!per) continue;
            tt = ((tt % nt) + nt) % nt;
        }
        bool dup = false;
        for (int q = 0; q < ns; ++q) dup = dup || (seen[q] == tt);
        if (dup) continue;
        seen[ns++] = tt;
        const int l0 = i - (tt << lt);
        const int nrep = per ? 3 : 1;
        for (int r = 0; r < nrep; ++r) {
            const int l = per ? l0 + (r - 1) * n : l0;
            if (l >= -1 && l <= t) {
                tl[m] = tt;
                el[m] = l + 1;
                ++m;
            }
        }
    }
    return m;
}

}  // namespace rh
