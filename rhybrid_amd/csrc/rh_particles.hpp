// rh_particles.hpp -- device kernels of the particle half of the step:
// Boris-Buneman push, the cell migration as a deterministic counting sort by
// subcell, the cloud-in-cell deposit with fixed-order reductions, the FP64
// cross-tile ghost fold, particle boundary conditions and FP64 diagnostics.
//
// Storage (per species, structure of arrays, Real = float or double):
//   x,y,z    cell-local position in units of dx, in [0,1); x = -1 marks a
//            particle removed by the boundary conditions
//   vx,vy,vz velocity in m/s
//   lc       the particle's cell inside its tile (implied by the sort order,
//            kept so that kernels need not search the subcell ranges)
// sorted by subcell key (rh_geom.hpp); sub[k] .. sub[k+1] is the range of
// subcell key k.
#pragma once

#include <hip/hip_runtime.h>

#include "rh_geom.hpp"

namespace rh {

template <typename Real>
struct PArrays {
    Real *x, *y, *z, *vx, *vy, *vz;
    uint8_t* lc;  // the particle's cell inside its tile (local index, < 64)
};

template <typename Real>
struct PushFields {
    const Real *fbx, *fby, *fbz;   // faceB components
    const Real *uex, *uey, *uez;   // cellUe
    const Real *epx, *epy, *epz;   // cellEp
};

struct PushParams {
    double dt, dt_dx;
    double qmdt2;        // 0.5 q dt / m
    double max_vi2;
    double GMdt;
    int accelerate, use_e_pressure, use_gravity, initial_flow_through;
    int still;           // sort only: no acceleration, no move
    uint8_t* live;       // pass 0 writes, per tile, whether it holds particles
};

enum DevCounter { C_MAX_VI = 0, C_MIN_CELL_RHO, C_MAX_CELL_UE, C_MIN_NODE_RHO, C_MAX_NODE_UE,
                  C_MAX_E, C_MAX_VW, C_DROPPED, C_DN_FALLBACK, C_RANK_OVERFLOW, C_N };

// Counters are spread over COUNTER_SLOTS copies (summed when read) so that
// waves of different blocks do not all hit one address.
constexpr int COUNTER_SLOTS = 64;

__device__ __forceinline__ void count_event(unsigned long long* counters, int id, bool pred) {
    // one atomic per wave: integer, so the total does not depend on order
    const unsigned long long b = __ballot(pred);
    const unsigned lane = (threadIdx.x + blockDim.x * threadIdx.y) & 63;
    if (b && lane == (unsigned)__builtin_ctzll(b)) {
        const unsigned slot = (blockIdx.x + 7u * blockIdx.y + 13u * blockIdx.z) & (COUNTER_SLOTS - 1);
        atomicAdd(counters + slot * C_N + id, (unsigned long long)__popcll(b));
    }
}

// The same for per-lane event counts, once per wave (every lane of the wave
// calls it)
__device__ __forceinline__ void count_events(unsigned long long* counters, int id, uint32_t n) {
    if (!__ballot(n != 0u)) return;
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if (((threadIdx.x + blockDim.x * threadIdx.y) & 63) == 0) {
        const unsigned slot = (blockIdx.x + 7u * blockIdx.y + 13u * blockIdx.z) & (COUNTER_SLOTS - 1);
        atomicAdd(counters + slot * C_N + id, (unsigned long long)n);
    }
}

// 2 / (1 + t^2) of the Boris rotation: the hardware reciprocal (1 ulp) in
// FP32, the correctly rounded quotient in FP64
__device__ __forceinline__ float boris_s_factor(float t2) { return 2.0f * __builtin_amdgcn_rcpf(1.0f + t2); }
__device__ __forceinline__ double boris_s_factor(double t2) { return 2.0 / (1.0 + t2); }

template <typename Real> __device__ __forceinline__ Real below_one();
template <> __device__ __forceinline__ float below_one<float>() { return 0x1.fffffep-1f; }
template <> __device__ __forceinline__ double below_one<double>() { return 0x1.fffffffffffffp-1; }

// The push and the cell migration, one workgroup of four waves per old tile,
// run twice with the same compiled body so that both passes compute bit-for-
// bit the same pushed state, destination slot and rank:
//   pass 0: BorisBuneman::propagate (src/particle_propagator_boris_buneman.h:
//           64-194, fields by face2r/cell2r, src/hybrid_propagator.cpp:
//           1740-1891), the move in FP64, and the destination slot in the
//           tile's extended region; counts particles per slot;
//   pass 1: the same, then writes each particle to (destination subcell's
//           start + the slot's offset inside it + rank).
// The tile's particles are taken 256 at a time.  Pass 0 counts slots with
// integer LDS atomics.  In pass 1 the rank of a particle is its position among
// the tile's particles with the same slot in memory order (the migration is a
// stable counting sort): each wave ranks its lanes with ballots and publishes
// its per-slot group sizes; a lane's destination is the slot's running
// counter (started, once per tile, at the slot's first destination index)
// plus its rank in the wave plus the group sizes of the earlier waves.
// The tile's field block is copied to LDS; removed particles carry x < 0.
constexpr int PM_THREADS = 256;
constexpr int PM_WAVES = PM_THREADS / 64;

RH_HD size_t pm_align8(size_t b) { return (b + 7) & ~size_t(7); }

// A tile's push fields as one contiguous block (32-bit words): faceB of the
// (t+1)^3 cells -1..t-1 (halo on the low side), cellUe and cellEp of the
// tile's cells, and the cells' neighbourhood masks.  Written once per field
// update by k_push_fields and copied to LDS by both passes of k_push_move.
template <typename Real>
RH_HD size_t push_field_words(const Geom& g) {
    const size_t nh = (size_t)(g.tx + 1) * (g.ty + 1) * (g.tz + 1);
    const size_t words = (sizeof(Real) * (3 * nh + 6 * (size_t)g.tc) + 4 * (size_t)g.tc) / 4;
    return (words + 3) & ~size_t(3);  // 16-byte blocks keep every tile's block aligned
}

template <typename Real>
RH_HD size_t push_move_lds_bytes(const Geom& g) {
    return pm_align8(4 * (size_t)g.extk + (size_t)PM_WAVES * g.extk) + 4 * push_field_words<Real>(g);
}

// Neighbourhood mask of a cell: bit (a+1) + 3(b+1) + 9(c+1) is set when the
// neighbour at offset (a,b,c) exists; bit 27 when all 26 exist
// (pargrid::ALL_NEIGHBOURS_EXIST).
constexpr int NMASK_ALL = 27;
RH_HD uint32_t neighbour_mask(const Geom& g, int i, int j, int k) {
    // existence is separable: offset (a,b,c) exists iff each axis offset does
    uint32_t ax[3] = {0, 0, 0};
    const int idx[3] = {i, j, k}, n[3] = {g.nx, g.ny, g.nz};
    for (int d = 0; d < 3; ++d)
        for (int a = -1; a <= 1; ++a)
            if (g.per[d] || (idx[d] + a >= 0 && idx[d] + a < n[d])) ax[d] |= 1u << (a + 1);
    uint32_t m = 0;
    for (int c = 0; c < 3; ++c)
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a)
                if ((ax[0] >> a) & (ax[1] >> b) & (ax[2] >> c) & 1u) m |= 1u << (a + 3 * b + 9 * c);
    if (ax[0] == 7u && ax[1] == 7u && ax[2] == 7u) m |= 1u << NMASK_ALL;
    return m;
}

// Fields of the push in LDS order (see push_field_words).  One thread per
// grid cell (coalesced reads of the nine arrays): the cell writes its faceB,
// cellUe, cellEp and neighbourhood mask into its own tile's block, and its
// faceB into the low-side halo of the tiles above it when it lies on its
// tile's upper faces (up to seven, periodic wrap included).  Halo entries of
// cells that do not exist are never written and keep the zero of the
// allocation.
template <typename Real>
__device__ __forceinline__ void push_fields_store(const Geom& g, int i, int j, int k, Real bx, Real by, Real bz,
                                                  Real uex, Real uey, Real uez, Real epx, Real epy, Real epz,
                                                  int use_e_pressure, uint32_t* __restrict__ out,
                                                  bool write_mask = true, bool nowrap = false) {
    const int hx = g.tx + 1, hy = g.ty + 1, nh = (g.tx + 1) * (g.ty + 1) * (g.tz + 1);
    const size_t nw = push_field_words<Real>(g);
    const int li = i & (g.tx - 1), lj = j & (g.ty - 1), lk = k & (g.tz - 1);
    const int lc = li + (lj << g.ltx) + (lk << (g.ltx + g.lty));
    {
        Real* blk = reinterpret_cast<Real*>(out + (size_t)tile_of(g, i, j, k) * nw);
        const int h = (li + 1) + hx * ((lj + 1) + hy * (lk + 1));
        blk[h] = bx; blk[nh + h] = by; blk[2 * nh + h] = bz;
        Real* sue = blk + 3 * nh;
        Real* sep = sue + 3 * g.tc;
        sue[lc] = uex; sue[g.tc + lc] = uey; sue[2 * g.tc + lc] = uez;
        if (use_e_pressure) { sep[lc] = epx; sep[g.tc + lc] = epy; sep[2 * g.tc + lc] = epz; }
        else { sep[lc] = sep[g.tc + lc] = sep[2 * g.tc + lc] = Real(0); }
        // (the mask depends on the grid only: a caller that rewrites the
        // fields of blocks written once in full may leave it)
        if (write_mask) reinterpret_cast<uint32_t*>(sep + 3 * g.tc)[lc] = neighbour_mask(g, i, j, k);
    }
    // halo copies: the cell is local coordinate -1 of the tile above along
    // each axis where it is the tile's last cell (nowrap: the caller knows
    // that those tiles lie inside the grid without a wrap)
    const int ux = li == g.tx - 1, uy = lj == g.ty - 1, uz = lk == g.tz - 1;
    for (int c = 0; c <= uz; ++c)
        for (int b = 0; b <= uy; ++b)
            for (int a = 0; a <= ux; ++a) {
                if (a + b + c == 0) continue;
                int ti = i + a, tj = j + b, tk = k + c;
                if (!nowrap) {
                    ti = wrap_axis(ti, g.nx, g.per[0]); tj = wrap_axis(tj, g.ny, g.per[1]); tk = wrap_axis(tk, g.nz, g.per[2]);
                    if (ti < 0 || tj < 0 || tk < 0) continue;
                }
                Real* blk = reinterpret_cast<Real*>(out + (size_t)tile_of(g, ti, tj, tk) * nw);
                const int h = (a ? 0 : li + 1) + hx * ((b ? 0 : lj + 1) + hy * (c ? 0 : lk + 1));
                blk[h] = bx; blk[nh + h] = by; blk[2 * nh + h] = bz;
            }
}

template <typename Real>
__device__ __forceinline__ void push_fields_at(const Geom& g, int i, int j, int k, int n, const PushFields<Real>& f,
                                               int use_e_pressure, uint32_t* __restrict__ out) {
    push_fields_store(g, i, j, k, f.fbx[n], f.fby[n], f.fbz[n], f.uex[n], f.uey[n], f.uez[n],
                      use_e_pressure ? f.epx[n] : Real(0), use_e_pressure ? f.epy[n] : Real(0),
                      use_e_pressure ? f.epz[n] : Real(0), use_e_pressure, out);
}

template <typename Real>
__global__ __launch_bounds__(256) void k_push_fields(Geom g, PushFields<Real> f, int use_e_pressure,
                                                     uint32_t* __restrict__ out) {
    const int i = blockIdx.x * 32 + threadIdx.x, j = blockIdx.y * 8 + threadIdx.y, k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    push_fields_at(g, i, j, k, cell_index(g, i, j, k), f, use_e_pressure, out);
}

// Injected particles waiting for their first push: per tile, the range of
// listed injection cells (cbeg, ntiles + 1 entries, null when none wait) and
// per listed cell the offset of its particles past the sorted store's end.
struct InjAppendix {
    const uint32_t* cbeg;
    const uint32_t* off;
};

template <typename Real>
__global__ __launch_bounds__(PM_THREADS) void k_push_move(Geom g, PushParams pp, const uint32_t* __restrict__ pfb,
                                                         PArrays<Real> A, PArrays<Real> D,
                                                         const uint32_t* __restrict__ sub, InjAppendix app,
                                                         uint16_t* __restrict__ cnt_tile,
                                                         const uint32_t* __restrict__ key_base,
                                                         unsigned long long* __restrict__ counters, int kbits,
                                                         int pass, uint32_t cap) {
    extern __shared__ __align__(16) unsigned char pm_smem[];
    const int NS = g.tc * 8;
    const int hx = g.tx + 1, hy = g.ty + 1, nh = (g.tx + 1) * (g.ty + 1) * (g.tz + 1);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(pm_smem);     // extk running slot counts
    uint8_t* grp = reinterpret_cast<uint8_t*>(cnt + g.extk);   // PM_WAVES x extk group sizes
    Real* fl = reinterpret_cast<Real*>(pm_smem + pm_align8(4 * (size_t)g.extk + (size_t)PM_WAVES * g.extk));
    Real* sbx = fl;            // faceB of local cells -1..t-1 (halo on the low side)
    Real* sby = fl + nh;
    Real* sbz = fl + 2 * nh;
    Real* sue = fl + 3 * nh;   // cellUe, cellEp of the tile's cells
    Real* sep = sue + 3 * g.tc;
    uint32_t* nmask = reinterpret_cast<uint32_t*>(sep + 3 * g.tc);  // neighbourhood of each tile cell
    const int T = pass == 1 ? xcd_brick_tile(g, (int)blockIdx.x) : xcd_tile(g, (int)blockIdx.x);
    if (T >= g.ntiles) return;
    int bx, by, bz;
    tile_coords(g, T, bx, by, bz);
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int ox = bx << g.ltx, oy = by << g.lty, oz = bz << g.ltz;
    // the tile's particle range first: its loads overlap the set-up below
    const uint32_t p0 = sub[(size_t)T * NS], p1 = sub[(size_t)(T + 1) * NS];
    // the tile's injected particles (past the store's end) follow its own
    uint32_t a0 = 0, na = 0;
    if (app.cbeg) {
        const uint32_t c0 = app.cbeg[T], c1 = app.cbeg[T + 1];
        if (c1 > c0) { a0 = sub[g.nkeys] + app.off[c0]; na = app.off[c1] - app.off[c0]; }
    }
    const uint32_t n0 = p1 - p0, ntot = n0 + na;
    if (pass == 0 && tid == 0) pp.live[T] = ntot > 0 ? 1 : 0;
    // an empty tile (common for a species confined to part of the grid):
    // nothing to push or move, and its slot table is not written (the slot
    // kernel reads the tables of live tiles only)
    if (ntot == 0) return;
    const uint32_t deadkey = (uint32_t)MAX_EXTK;
    // prefetch of the first chunk; an injected particle beyond the store's
    // capacity was never written (the injection counted the overflow)
    // (the tile's particles in chunks of PM_THREADS, prefetched two chunks
    // ahead into two register buffers used in turn: no copies between them,
    // so a buffer's loads are waited for only where its chunk is used)
    struct PBuf { Real X, Y, Z, vx, vy, vz; int lc; };
    auto load_chunk = [&](PBuf& b, uint32_t u) {
        const uint32_t p = u < n0 ? p0 + u : a0 + (u - n0);
        if (u < ntot) {
            if (p < cap) { b.X = A.x[p]; b.Y = A.y[p]; b.Z = A.z[p]; b.vx = A.vx[p]; b.vy = A.vy[p]; b.vz = A.vz[p]; b.lc = A.lc[p]; }
            else b.X = Real(-1);
        }
    };
    PBuf bufA{Real(0), Real(0), Real(0), Real(0), Real(0), Real(0), 0};
    PBuf bufB = bufA;
    load_chunk(bufA, tid);
    load_chunk(bufB, tid + PM_THREADS);
    // slot counters: the 8 octant slots of an extended cell are 8 consecutive
    // entries (16-byte vectors of counts, 32-byte vectors of subcell starts)
    uint4* cnt4 = reinterpret_cast<uint4*>(cnt);
    if (pass == 0) {
        for (int e = tid; e < g.extk / 4; e += PM_THREADS) cnt4[e] = make_uint4(0u, 0u, 0u, 0u);
    } else {
        // each slot's first destination index: the destination subcell's
        // start (key_base) plus the slot's offset inside it (cnt_tile); the
        // running counter then grows by the slot's particles in order
        for (int ec = tid; ec < g.next; ec += PM_THREADS) {
            const int lx = ec % g.ex - 1, ly = (ec / g.ex) % g.ey - 1, lz = ec / (g.ex * g.ey) - 1;
            int tX = bx + (lx < 0 ? -1 : (lx >= g.tx ? 1 : 0));
            int tY = by + (ly < 0 ? -1 : (ly >= g.ty ? 1 : 0));
            int tZ = bz + (lz < 0 ? -1 : (lz >= g.tz ? 1 : 0));
            // (outside a non-periodic grid the slot is never used; the wrap
            // keeps the index in range)
            tX += tX < 0 ? g.ntx : (tX >= g.ntx ? -g.ntx : 0);
            tY += tY < 0 ? g.nty : (tY >= g.nty ? -g.nty : 0);
            tZ += tZ < 0 ? g.ntz : (tZ >= g.ntz ? -g.ntz : 0);
            const uint32_t nl = (uint32_t)((lx & (g.tx - 1)) + ((ly & (g.ty - 1)) << g.ltx) +
                                           ((lz & (g.tz - 1)) << (g.ltx + g.lty)));
            const uint32_t gk = ((uint32_t)(tX + g.ntx * (tY + g.nty * tZ)) * (uint32_t)g.tc + nl) * 8u;
            const uint4 b0 = *reinterpret_cast<const uint4*>(key_base + gk);
            const uint4 b1 = *reinterpret_cast<const uint4*>(key_base + gk + 4);
            const uint4 c = *reinterpret_cast<const uint4*>(cnt_tile + ((size_t)T * g.next + ec) * 8);
            cnt4[2 * ec] = make_uint4(b0.x + (c.x & 0xFFFFu), b0.y + (c.x >> 16), b0.z + (c.y & 0xFFFFu), b0.w + (c.y >> 16));
            cnt4[2 * ec + 1] = make_uint4(b1.x + (c.z & 0xFFFFu), b1.y + (c.z >> 16), b1.z + (c.w & 0xFFFFu), b1.w + (c.w >> 16));
        }
    }
    for (int e = tid; e < PM_WAVES * g.extk / 16; e += PM_THREADS) reinterpret_cast<uint4*>(grp)[e] = make_uint4(0u, 0u, 0u, 0u);
    {
        const int nw4 = (int)push_field_words<Real>(g) / 4;
        const uint4* src = reinterpret_cast<const uint4*>(pfb) + (size_t)T * nw4;
        uint4* dst = reinterpret_cast<uint4*>(fl);
        for (int q = tid; q < nw4; q += PM_THREADS) dst[q] = src[q];
    }
    __syncthreads();
    uint32_t n_clamped = 0, n_dropped = 0;  // this lane's events (pass 0)
    auto chunk = [&](uint32_t ub, PBuf& buf) {
        const uint32_t u = ub + tid;
        const bool act = u < ntot;
        Real X = buf.X, Y = buf.Y, Z = buf.Z, vx = buf.vx, vy = buf.vy, vz = buf.vz;
        const int lc = buf.lc;
        load_chunk(buf, u + 2 * PM_THREADS);
        uint32_t key = deadkey;
        bool clamped = false, dropped = false;
        int di = 0, dj = 0, dk = 0;  // destination cell, local coordinates of the source tile
        if (act && X >= Real(0)) {
            int li, lj, lk;
            local_ijk(g, lc, li, lj, lk);
            const uint32_t nm = nmask[lc];
            const bool accel = pp.accelerate && ((nm >> NMASK_ALL) & 1u) && !pp.initial_flow_through;
            if (accel) {
                const int h = (li + 1) + hx * ((lj + 1) + hy * (lk + 1));
                // face2r: each component linear along its own axis only
                const Real Bx = X * sbx[h] + (Real(1) - X) * sbx[h - 1];
                const Real By = Y * sby[h] + (Real(1) - Y) * sby[h - hx];
                const Real Bz = Z * sbz[h] + (Real(1) - Z) * sbz[h - hx * hy];
                // cell2r: nearest grid point; E = B x Ue (= -Ue x B) (+ Ep)
                const Real Ux = sue[lc], Uy = sue[g.tc + lc], Uz = sue[2 * g.tc + lc];
                Real Ex = By * Uz - Bz * Uy;
                Real Ey = Bz * Ux - Bx * Uz;
                Real Ez = Bx * Uy - By * Ux;
                Ex += sep[lc]; Ey += sep[g.tc + lc]; Ez += sep[2 * g.tc + lc];
                const Real qm = (Real)pp.qmdt2;
                const Real dvx = qm * Ex, dvy = qm * Ey, dvz = qm * Ez;
                const Real tx = qm * Bx, ty = qm * By, tz = qm * Bz;
                const Real t2 = tx * tx + ty * ty + tz * tz;
                const Real b2 = boris_s_factor(t2);
                const Real sx = b2 * tx, sy = b2 * ty, sz = b2 * tz;
                const Real vmx = vx + dvx, vmy = vy + dvy, vmz = vz + dvz;
                const Real v0x = vmx + vmy * tz - vmz * ty;
                const Real v0y = vmy + vmz * tx - vmx * tz;
                const Real v0z = vmz + vmx * ty - vmy * tx;
                vx = vmx + v0y * sz - v0z * sy + dvx;
                vy = vmy + v0z * sx - v0x * sz + dvy;
                vz = vmz + v0x * sy - v0y * sx + dvz;
                if (pp.use_gravity) {
                    const double rx = g.x0 + (ox + li + (double)X) * g.dx;
                    const double ry = g.y0 + (oy + lj + (double)Y) * g.dx;
                    const double rz = g.z0 + (oz + lk + (double)Z) * g.dx;
                    const double r = sqrt(rx * rx + ry * ry + rz * rz);
                    const double r3 = r * r * r;
                    if (r3 > 0) {
                        const double s = -pp.GMdt / r3;
                        vx += (Real)(s * rx); vy += (Real)(s * ry); vz += (Real)(s * rz);
                    }
                }
                const Real v2 = vx * vx + vy * vy + vz * vz;
                if (v2 > (Real)pp.max_vi2) {
                    const Real norm = sqrt((Real)pp.max_vi2 / v2);
                    vx *= norm; vy *= norm; vz *= norm;
                    clamped = true;
                }
            }
            // move in FP64; the integer part is the cell crossing (a still
            // pass keeps every particle where it is)
            int sx = 0, sy = 0, sz = 0;
            bool ok = true;
            if (!pp.still) {
                const double Xd = (double)X + (double)vx * pp.dt_dx;
                const double Yd = (double)Y + (double)vy * pp.dt_dx;
                const double Zd = (double)Z + (double)vz * pp.dt_dx;
                const double fx = floor(Xd), fy = floor(Yd), fz = floor(Zd);
                ok = (fx >= -1.0 && fx <= 1.0 && fy >= -1.0 && fy <= 1.0 && fz >= -1.0 && fz <= 1.0);
                sx = (int)fx; sy = (int)fy; sz = (int)fz;
                X = (Real)(Xd - fx); Y = (Real)(Yd - fy); Z = (Real)(Zd - fz);
                if (X >= Real(1)) X = below_one<Real>();
                if (Y >= Real(1)) Y = below_one<Real>();
                if (Z >= Real(1)) Z = below_one<Real>();
            }
            if (ok && ((nm >> ((sx + 1) + 3 * (sy + 1) + 9 * (sz + 1))) & 1u)) {
                const int o = (X >= Real(0.5)) + 2 * (Y >= Real(0.5)) + 4 * (Z >= Real(0.5));
                di = li + sx; dj = lj + sy; dk = lk + sz;
                key = (uint32_t)(((di + 1) + g.ex * ((dj + 1) + g.ey * (dk + 1))) * 8 + o);
            } else {
                dropped = true;
            }
        }
        if (pass == 0) {
            // counts only: integer LDS atomics, so the order does not matter;
            // the events are counted once per tile
            if (key != deadkey) atomicAdd(&cnt[key], 1u);
            n_clamped += clamped;
            n_dropped += dropped;
            return;
        }
        // tiles divide the grid, so a periodic wrap keeps the local index
        const uint8_t nlcell = (uint8_t)((di & (g.tx - 1)) + ((dj & (g.ty - 1)) << g.ltx) +
                                         ((dk & (g.tz - 1)) << (g.ltx + g.lty)));
        // rank among this wave's lanes holding the same slot
        unsigned long long match = __ballot(key != deadkey);
        for (int b = 0; b < kbits; ++b) {
            const bool bit = (key >> b) & 1u;
            const unsigned long long bal = __ballot(bit);
            match &= bit ? bal : ~bal;
        }
        const unsigned long long below = (lane == 0) ? 0ull : (match & ((~0ull) >> (64 - lane)));
        const uint32_t rin = (uint32_t)__popcll(below);
        const uint32_t gsz = (uint32_t)__popcll(match);
        const bool leader = key != deadkey && rin == gsz - 1;
        if (leader) grp[w * g.extk + key] = (uint8_t)gsz;
        __syncthreads();
        if (key != deadkey) {
            uint32_t dst = cnt[key] + rin;  // the slot's next destination index
            for (int v = 0; v < w; ++v) dst += grp[v * g.extk + key];
            if (dst < cap) {
                D.x[dst] = X; D.y[dst] = Y; D.z[dst] = Z;
                D.vx[dst] = vx; D.vy[dst] = vy; D.vz[dst] = vz;
                D.lc[dst] = nlcell;
            } else {
                atomicAdd(counters + C_RANK_OVERFLOW, 1ull);  // slot 0
            }
        }
        __syncthreads();
        if (leader) {
            atomicAdd(&cnt[key], gsz);  // integer: the order of the four waves does not matter
            grp[w * g.extk + key] = 0;
        }
        __syncthreads();
    };
    for (uint32_t ub = 0; ub < ntot; ub += 2 * PM_THREADS) {
        chunk(ub, bufA);
        if (ub + PM_THREADS < ntot) chunk(ub + PM_THREADS, bufB);  // the same for the whole workgroup
    }
    __syncthreads();
    if (pass == 0) {
        count_events(counters, C_MAX_VI, n_clamped);
        count_events(counters, C_DROPPED, n_dropped);
    }
    if (pass == 0)
        for (int ec = tid; ec < g.next; ec += PM_THREADS) {
            const uint4 a = cnt4[2 * ec], b = cnt4[2 * ec + 1];
            const uint32_t hi = (a.x | a.y | a.z | a.w | b.x | b.y | b.z | b.w) >> 16;
            if (hi) atomicAdd(counters + C_RANK_OVERFLOW, 1ull);  // slot 0
            *reinterpret_cast<uint4*>(cnt_tile + ((size_t)T * g.next + ec) * 8) =
                make_uint4((a.x & 0xFFFFu) | (a.y << 16), (a.z & 0xFFFFu) | (a.w << 16),
                           (b.x & 0xFFFFu) | (b.y << 16), (b.z & 0xFFFFu) | (b.w << 16));
        }
}

// Number of particles arriving in each subcell: the slot counts of every tile
// whose extended region holds the cell, in a fixed order (out = totals, to be
// scanned into the subcells' first destination index); the slot counts are
// replaced by each slot's offset inside its subcell's range.  One thread per
// cell: a cell's eight octant slots are eight consecutive 16-bit entries of
// every candidate tile's table, read and written as one 16-byte vector.
// axis_candidates for an axis with at least three tiles or without wrap: the
// previous tile (cell on the tile's low face), the own tile, the next tile
// (cell on the high face), in that order.
__device__ __forceinline__ int axis_candidates_fast(int i, int lt, int nt, int per, int* tl, int* el) {
    const int t = 1 << lt, t0 = i >> lt, l = i & (t - 1);
    int m = 0;
    if (l == 0 && (t0 > 0 || per)) { tl[m] = t0 > 0 ? t0 - 1 : nt - 1; el[m++] = t + 1; }
    tl[m] = t0; el[m++] = l + 1;
    if (l == t - 1 && (t0 < nt - 1 || per)) { tl[m] = t0 < nt - 1 ? t0 + 1 : 0; el[m++] = 0; }
    return m;
}

// The slot offsets of one candidate (eight packed 16-bit counts) before its
// counts are added to the running totals.
__device__ __forceinline__ uint4 slot_offsets8(uint32_t (&run)[8], uint4 v) {
    const uint4 off = make_uint4((run[0] & 0xFFFFu) | (run[1] << 16), (run[2] & 0xFFFFu) | (run[3] << 16),
                                 (run[4] & 0xFFFFu) | (run[5] << 16), (run[6] & 0xFFFFu) | (run[7] << 16));
    run[0] += v.x & 0xFFFFu; run[1] += v.x >> 16; run[2] += v.y & 0xFFFFu; run[3] += v.y >> 16;
    run[4] += v.z & 0xFFFFu; run[5] += v.z >> 16; run[6] += v.w & 0xFFFFu; run[7] += v.w >> 16;
    return off;
}

constexpr int SLOT_THREADS = 256;

__global__ __launch_bounds__(SLOT_THREADS) void k_slot_offsets(Geom g, uint16_t* __restrict__ cnt_tile,
                                                              const uint8_t* __restrict__ live,
                                                              uint32_t* __restrict__ out,
                                                              unsigned long long* __restrict__ counters) {
    // blocks dealt to the XCDs in contiguous ranges (slab order): a tile's
    // slot table is read by its own cells and its neighbours' face cells
    const uint32_t nb = ((uint32_t)g.ntiles * (uint32_t)g.tc + SLOT_THREADS - 1) / SLOT_THREADS;
    const uint32_t blk = (uint32_t)xcd_index((int)nb, (int)blockIdx.x);
    if (blk >= nb) return;
    const uint32_t id = blk * (uint32_t)SLOT_THREADS + threadIdx.x;  // = tile * tc + local cell
    if (id >= (uint32_t)g.ntiles * (uint32_t)g.tc) return;
    const int T = (int)(id >> (g.ltx + g.lty + g.ltz)), lc = (int)(id & (uint32_t)(g.tc - 1));
    int li, lj, lk;
    local_ijk(g, lc, li, lj, lk);
    const int i = ((T % g.ntx) << g.ltx) + li, j = (((T / g.ntx) % g.nty) << g.lty) + lj,
              k = ((T / (g.ntx * g.nty)) << g.ltz) + lk;
    int tx[6], ex[6], ty[6], ey[6], tz[6], ez[6];
    const bool fx = !g.per[0] || g.ntx >= 3, fy = !g.per[1] || g.nty >= 3, fz = !g.per[2] || g.ntz >= 3;
    uint4* tab = reinterpret_cast<uint4*>(cnt_tile);  // [tile][extended cell]: eight octant slots
    uint32_t run[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
    if (fx && fy && fz && g.tx >= 2 && g.ty >= 2 && g.tz >= 2) {
        // tile sides >= 2: a cell is on at most one face of its tile per
        // axis, so at most two candidates per axis; all eight loads are
        // issued before the ordered accumulation (32-bit slot offsets: the
        // context rejects tilings whose slot table has 2^32 entries or more)
        const int mx = axis_candidates_fast(i, g.ltx, g.ntx, g.per[0], tx, ex);
        const int my = axis_candidates_fast(j, g.lty, g.nty, g.per[1], ty, ey);
        const int mz = axis_candidates_fast(k, g.ltz, g.ntz, g.per[2], tz, ez);
        // (an empty candidate tile has no slot table: it adds nothing)
        uint32_t slot[8];
        uint4 v[8];
        bool use[8];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    const int q = a + 2 * b + 4 * c;
                    const bool inb = a < mx && b < my && c < mz;
                    const int TT = inb ? tx[a] + g.ntx * (ty[b] + g.nty * tz[c]) : T;
                    use[q] = inb && live[TT];
                    slot[q] = (uint32_t)TT * (uint32_t)g.next + (uint32_t)(inb ? ex[a] + g.ex * (ey[b] + g.ey * ez[c]) : 0);
                    v[q] = use[q] ? tab[slot[q]] : make_uint4(0u, 0u, 0u, 0u);
                }
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (use[q]) tab[slot[q]] = slot_offsets8(run, v[q]);  // offsets inside the subcells' ranges
    } else {
        const int mx = fx ? axis_candidates_fast(i, g.ltx, g.ntx, g.per[0], tx, ex)
                          : axis_candidates(i, g.nx, g.ltx, g.ntx, g.per[0], tx, ex);
        const int my = fy ? axis_candidates_fast(j, g.lty, g.nty, g.per[1], ty, ey)
                          : axis_candidates(j, g.ny, g.lty, g.nty, g.per[1], ty, ey);
        const int mz = fz ? axis_candidates_fast(k, g.ltz, g.ntz, g.per[2], tz, ez)
                          : axis_candidates(k, g.nz, g.ltz, g.ntz, g.per[2], tz, ez);
        for (int c = 0; c < mz; ++c)
            for (int b = 0; b < my; ++b)
                for (int a = 0; a < mx; ++a) {
                    const int TT = tx[a] + g.ntx * (ty[b] + g.nty * tz[c]);
                    if (!live[TT]) continue;
                    const size_t slot = (size_t)TT * g.next + (size_t)(ex[a] + g.ex * (ey[b] + g.ey * ez[c]));
                    tab[slot] = slot_offsets8(run, tab[slot]);
                }
    }
    uint4* o4 = reinterpret_cast<uint4*>(out + (size_t)id * 8);
    o4[0] = make_uint4(run[0], run[1], run[2], run[3]);
    o4[1] = make_uint4(run[4], run[5], run[6], run[7]);
    if ((run[0] | run[1] | run[2] | run[3] | run[4] | run[5] | run[6] | run[7]) > 0xFFFFu)
        atomicAdd(counters + C_RANK_OVERFLOW, 1ull);  // offsets are 16-bit
}

// Exclusive scan of uint32 counts (three passes: block sums, scan of the block
// sums, block scan plus offset).  SCAN_ITEMS elements per thread.
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* smem, uint32_t& total) {
    const int t = threadIdx.x;
    smem[t] = v;
    __syncthreads();
    for (int off = 1; off < SCAN_THREADS; off <<= 1) {
        const uint32_t a = (t >= off) ? smem[t - off] : 0u;
        __syncthreads();
        smem[t] += a;
        __syncthreads();
    }
    total = smem[SCAN_THREADS - 1];
    const uint32_t incl = smem[t];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint32_t* __restrict__ in, uint32_t n,
                                                             uint32_t* __restrict__ bsum) {
    __shared__ uint32_t smem[SCAN_THREADS];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
    for (int q = 0; q < SCAN_ITEMS; ++q) {  // coalesced: item q of thread t at base + q*threads + t
        const size_t i = base + (size_t)q * SCAN_THREADS + threadIdx.x;
        if (i < n) s += in[i];
    }
    uint32_t tot;
    block_exclusive_scan(s, smem, tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// In-place safe: a block reads its whole tile into LDS (coalesced) before any
// write.  LDS element j sits at j + j/SCAN_ITEMS, so that the per-thread runs
// of SCAN_ITEMS consecutive elements do not share banks.
__device__ __forceinline__ int scan_pad(int j) { return j + j / SCAN_ITEMS; }

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const uint32_t* __restrict__ in, uint32_t n,
                                                            const uint32_t* __restrict__ boff,
                                                            uint32_t* __restrict__ out) {
    __shared__ uint32_t smem[SCAN_THREADS];
    __shared__ uint32_t tile[SCAN_TILE + SCAN_THREADS];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE;
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        const size_t i = base + (size_t)q * SCAN_THREADS + threadIdx.x;
        tile[scan_pad(q * SCAN_THREADS + threadIdx.x)] = i < n ? in[i] : 0u;
    }
    __syncthreads();
    const int t0 = threadIdx.x * (SCAN_ITEMS + 1);
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        v[q] = tile[t0 + q];
        s += v[q];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, smem, tot) + boff[blockIdx.x];
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        tile[t0 + q] = run;
        run += v[q];
    }
    __syncthreads();
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        const size_t i = base + (size_t)q * SCAN_THREADS + threadIdx.x;
        if (i < n) out[i] = tile[scan_pad(q * SCAN_THREADS + threadIdx.x)];
    }
    // the grand total goes one past the end
    if (base + SCAN_TILE >= n && threadIdx.x == SCAN_THREADS - 1) out[n] = boff[blockIdx.x] + tot;
}

// Cloud-in-cell deposit (Accumulator::accumulateCell,
// src/particle_accumulator.cpp:43-156).  One workgroup per tile, one thread
// per subcell: a subcell's particles share one 2x2x2 footprint, so the thread
// sums, in particle order, the 8 corner weights times (1, vx, vy, vz).  The
// eight octant threads of a cell then combine their corner sums into the
// cell's 27 targets by lane exchanges, the targets are staged in LDS, and
// every cell of the tile's extended region adds the (up to 27) cell terms
// that reach it in a fixed order (crowded tiles: eight threads per subcell
// and the (up to 64) subcell terms, two quantities at a time).  The result
// is the tile's ghost partial of this species (unscaled); k_fold sums
// partials over tiles and species in FP64.  No atomics, so every run gives
// the same bits.
template <typename Real>
__device__ __forceinline__ void cic_accumulate(Real (&acc)[4][8], bool hx, bool hy, bool hz, Real X, Real Y, Real Z,
                                               Real vx, Real vy, Real vz) {
    // weights of the lower / upper target along each axis
    const Real wx1 = hx ? X - Real(0.5) : X + Real(0.5);
    const Real wy1 = hy ? Y - Real(0.5) : Y + Real(0.5);
    const Real wz1 = hz ? Z - Real(0.5) : Z + Real(0.5);
    const Real live = X >= Real(0) ? Real(1) : Real(0);  // removed particles add nothing
    const Real wx[2] = {live - live * wx1, live * wx1}, wy[2] = {Real(1) - wy1, wy1}, wz[2] = {Real(1) - wz1, wz1};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const Real w = wx[c & 1] * wy[(c >> 1) & 1] * wz[c >> 2];
        acc[0][c] += w;
        acc[1][c] += w * vx;
        acc[2][c] += w * vy;
        acc[3][c] += w * vz;
    }
}

// Octant block of the deposit's LDS staging: the tile's cells padded by 8
// words, so that the 8 octants x 8 cells a wave writes fall in distinct banks
// (octant stride = 8 banks).
RH_HD int deposit_octant_block(const Geom& g) { return g.tc + 8; }

// Two of the four quantities (the charge and vx, or vy and vz): a and b
// are the particle's factors for them (1 and vx, or vy and vz).
template <typename Real>
__device__ __forceinline__ void cic_accumulate2(Real (&acc)[2][8], bool hx, bool hy, bool hz, Real X, Real Y,
                                                Real Z, Real a, Real b) {
    const Real wx1 = hx ? X - Real(0.5) : X + Real(0.5);
    const Real wy1 = hy ? Y - Real(0.5) : Y + Real(0.5);
    const Real wz1 = hz ? Z - Real(0.5) : Z + Real(0.5);
    const Real live = X >= Real(0) ? Real(1) : Real(0);
    const Real wx[2] = {live - live * wx1, live * wx1}, wy[2] = {Real(1) - wy1, wy1}, wz[2] = {Real(1) - wz1, wz1};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const Real w = wx[c & 1] * wy[(c >> 1) & 1] * wz[c >> 2];
        acc[0][c] += w * a;
        acc[1][c] += w * b;
    }
}

// LDS of the deposit: half of the subcell sums (2 quantities x 8 corners) as
// [quantity*8+corner][octant][cell] (octant blocks of deposit_octant_block
// words); the first half's results.
template <typename Real>
RH_HD size_t deposit_lds_bytes(const Geom& g) {
    return sizeof(Real) * (16 * 8 * (size_t)deposit_octant_block(g) + 2 * (size_t)g.next);
}

// For every cell of the extended region, the LDS offsets (corner plane,
// octant block, cell) of the subcell sums that reach it: cell + half - 1 +
// corner = target along each axis.  The same for every tile of a shape;
// built once on the host, at most 64 entries per cell.
inline int deposit_gather_lists(const Geom& g, uint16_t* list, uint8_t* count) {
    const int OB = deposit_octant_block(g), QS = 8 * OB;
    for (int e = 0; e < g.next; ++e) {
        const int lx = e % g.ex - 1, ly = (e / g.ex) % g.ey - 1, lz = e / (g.ex * g.ey) - 1;
        int m = 0;
        for (int hz = 0; hz < 2; ++hz)
            for (int kz = 0; kz < 2; ++kz) {
                const int cz = lz + 1 - hz - kz;
                if (cz < 0 || cz >= g.tz) continue;
                for (int hy = 0; hy < 2; ++hy)
                    for (int ky = 0; ky < 2; ++ky) {
                        const int cy = ly + 1 - hy - ky;
                        if (cy < 0 || cy >= g.ty) continue;
                        for (int hx = 0; hx < 2; ++hx)
                            for (int kx = 0; kx < 2; ++kx) {
                                const int cx = lx + 1 - hx - kx;
                                if (cx < 0 || cx >= g.tx) continue;
                                const int cell = cx + g.tx * (cy + g.ty * cz);
                                const int corner = kx + 2 * ky + 4 * kz, oct = hx + 2 * hy + 4 * hz;
                                list[e * 64 + m++] = (uint16_t)(corner * QS + oct * OB + cell);
                            }
                    }
            }
        // ascending offsets: a fixed order that also walks LDS forward
        for (int a = 1; a < m; ++a)
            for (int b = a; b > 0 && list[e * 64 + b - 1] > list[e * 64 + b]; --b) {
                const uint16_t t = list[e * 64 + b];
                list[e * 64 + b] = list[e * 64 + b - 1];
                list[e * 64 + b - 1] = t;
            }
        count[e] = (uint8_t)m;
        // padded to a multiple of 8 entries with a word the deposit keeps at
        // zero (corner 0, octant 0, the cell past the tile's: cell padding)
        for (int t = m; t < ((m + 7) & ~7); ++t) list[e * 64 + t] = (uint16_t)g.tc;
    }
    return 0;
}

// The one-lane-per-subcell path first combines, in registers, the corner
// sums of a cell's eight octant lanes into the cell's 27 targets (cell + d,
// d in {-1, 0, 1}^3), then gathers for every cell of the extended region the
// (up to 27) cell terms that reach it.  LDS: [quantity][target][cell], cell
// blocks of deposit_octant_block words.  The lists hold, for every cell of
// the extended region, the offsets (target, cell) of its terms in ascending
// order, at most 27 per cell (stride 32), padded to a multiple of 8 entries
// with the offset of a word the deposit keeps at zero (target 0, the cell
// past the tile's: inside the padding of the cell blocks); built once on the
// host.
inline int deposit_target_lists(const Geom& g, uint16_t* list, uint8_t* count) {
    const int CB = deposit_octant_block(g);
    for (int e = 0; e < g.next; ++e) {
        const int lx = e % g.ex - 1, ly = (e / g.ex) % g.ey - 1, lz = e / (g.ex * g.ey) - 1;
        int m = 0;
        for (int dz = -1; dz <= 1; ++dz) {
            const int cz = lz - dz;
            if (cz < 0 || cz >= g.tz) continue;
            for (int dy = -1; dy <= 1; ++dy) {
                const int cy = ly - dy;
                if (cy < 0 || cy >= g.ty) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    const int cx = lx - dx;
                    if (cx < 0 || cx >= g.tx) continue;
                    const int t27 = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
                    list[e * 32 + m++] = (uint16_t)(t27 * CB + cx + g.tx * (cy + g.ty * cz));
                }
            }
        }
        for (int a = 1; a < m; ++a)
            for (int b = a; b > 0 && list[e * 32 + b - 1] > list[e * 32 + b]; --b) {
                const uint16_t t = list[e * 32 + b];
                list[e * 32 + b] = list[e * 32 + b - 1];
                list[e * 32 + b - 1] = t;
            }
        count[e] = (uint8_t)m;
        for (int t = m; t < ((m + 7) & ~7); ++t) list[e * 32 + t] = (uint16_t)g.tc;
    }
    return 0;
}

// The value of lane ^ M (M = 1, 2, 4: inside the 8 lanes of one cell's
// octants): DPP quad permutations for 1 and 2, a swizzle in xor mode for 4.
template <int M>
__device__ __forceinline__ int octant_xor_bits(int v) {
    if constexpr (M == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    else if constexpr (M == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    else return __builtin_amdgcn_ds_swizzle(v, (M << 10) | 0x1F);  // and 31, or 0, xor M
}
template <int M>
__device__ __forceinline__ float octant_xor(float v) { return __int_as_float(octant_xor_bits<M>(__float_as_int(v))); }
template <int M>
__device__ __forceinline__ double octant_xor(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = octant_xor_bits<M>((int)(b & 0xFFFFFFFFll)), hi = octant_xor_bits<M>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// One exchange stage along an axis with octant bit `bit` (lane mask M):
// pairs (c0, c1) of corner 0 and 1 along the axis.  The lane with bit 0 keeps
// target -1 (its corner 0) in c0 and target 0 (its corner 1 plus the
// partner's corner 0) in c1; the lane with bit 1 keeps target +1 (its corner
// 1) in c0.  A fixed order: the same bits in every run.
template <int M, typename Real>
__device__ __forceinline__ void octant_stage(Real& c0, Real& c1, bool bit) {
    const Real p = octant_xor<M>(c0);
    const Real a0 = c0, a1 = c1;
    c0 = bit ? a1 : a0;
    c1 = bit ? Real(0) : a1 + p;
}

// Particles per chunk when the one-lane-per-subcell path stages the tile's
// particles (6 arrays) through LDS: as many as fit in the subcell-sum
// staging, which the chunks alias (it is written only after the last one).
template <typename Real>
RH_HD int deposit_chunk(const Geom& g) {
    return ((16 * 8 * deposit_octant_block(g)) / 6) & ~63;
}

template <typename Real>
struct alignas(4 * sizeof(Real)) Real4 { Real v[4]; };
template <typename Real>
struct alignas(2 * sizeof(Real)) Real2 { Real v[2]; };

// The deposit of a crowded tile (k_deposit sends tiles with at least
// dense_min particles here): eight lanes share a subcell and read eight
// consecutive particles per array (one cache line per group instead of one
// per lane), then add their sums in a fixed butterfly order; blockDim/8
// subcells per round.  Within the same LDS as the sparse path, the two
// halves (charge and vx, then vy and vz) each read the particles once.
template <typename Real>
__device__ void deposit_dense_tile(const Geom& g, const PArrays<Real>& A, const uint32_t* __restrict__ sub,
                                   const uint16_t* __restrict__ glist, const uint8_t* __restrict__ gcount,
                                   Real* __restrict__ partial, int T, Real* S, Real* R) {
    const int NS = g.tc * 8;
    const int OB = deposit_octant_block(g);
    const int QS = 8 * OB;
    const int tid = threadIdx.x, grp = tid >> 3, gl = tid & 7, ngrp = (int)blockDim.x >> 3;
    Real4<Real>* out = reinterpret_cast<Real4<Real>*>(partial) + (size_t)T * g.next;
    for (int half = 0; half < 2; ++half) {
        if (half) __syncthreads();  // the first half's reads are done
        for (int base = 0; base < NS; base += ngrp) {
            const int ls = base + grp;
            const bool valid = ls < NS;
            const uint32_t key = (uint32_t)(T * NS + (valid ? ls : 0));
            const uint32_t p0 = valid ? sub[key] : 0u, p1 = valid ? sub[key + 1] : 0u;
            const int o = ls & 7, cell = ls >> 3;
            const bool hx = o & 1, hy = (o >> 1) & 1, hz = (o >> 2) & 1;
            Real acc[2][8];
#pragma unroll
            for (int q = 0; q < 2; ++q)
#pragma unroll
                for (int c = 0; c < 8; ++c) acc[q][c] = Real(0);
            for (uint32_t p = p0 + gl; p < p1; p += 8) {
                const Real a = half ? A.vy[p] : Real(1), b = half ? A.vz[p] : A.vx[p];
                cic_accumulate2(acc, hx, hy, hz, A.x[p], A.y[p], A.z[p], a, b);
            }
#pragma unroll
            for (int off = 1; off < 8; off <<= 1)
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                    for (int c = 0; c < 8; ++c) acc[q][c] += __shfl_xor(acc[q][c], off);
            if (valid && gl == 0)
#pragma unroll
                for (int q = 0; q < 2; ++q)
#pragma unroll
                    for (int c = 0; c < 8; ++c) S[(q * 8 + c) * QS + o * OB + cell] = acc[q][c];
        }
        if (tid < 2) S[tid * 8 * QS + g.tc] = Real(0);  // the lists' padding entries read these words
        __syncthreads();
        for (int e = tid; e < g.next; e += blockDim.x) {
            // eight entries at a time (the padding adds zeros after the terms)
            const int ng = (gcount[e] + 7) >> 3;
            const uint4* lst = reinterpret_cast<const uint4*>(glist + (size_t)e * 64);
            Real res[2] = {Real(0), Real(0)};
            for (int gi = 0; gi < ng; ++gi) {
                const uint4 l8 = lst[gi];
                const uint32_t w[4] = {l8.x, l8.y, l8.z, l8.w};
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const int off = (int)((w[t >> 1] >> (16 * (t & 1))) & 0xFFFFu);
                    res[0] += S[off];
                    res[1] += S[8 * QS + off];
                }
            }
            if (half == 0) {
                R[2 * e] = res[0];
                R[2 * e + 1] = res[1];
            } else {
                Real4<Real> r;
                r.v[0] = R[2 * e]; r.v[1] = R[2 * e + 1]; r.v[2] = res[0]; r.v[3] = res[1];
                out[e] = r;
            }
        }
    }
}

template <typename Real>
__global__ __launch_bounds__(512) void k_deposit(Geom g, PArrays<Real> A, const uint32_t* __restrict__ sub,
                                                 const uint16_t* __restrict__ glist,
                                                 const uint8_t* __restrict__ gcount,
                                                 const uint16_t* __restrict__ tlist,
                                                 const uint8_t* __restrict__ tcount, Real* __restrict__ partial,
                                                 uint32_t dense_min) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int NS = g.tc * 8;
    const int OB = deposit_octant_block(g);  // octant block (padded)
    const int QS = 8 * OB;            // one (quantity, corner) plane
    Real* S = reinterpret_cast<Real*>(smem_raw);
    Real* R = S + 16 * QS;  // first-half results, next x 2
    const int T = xcd_tile(g, (int)blockIdx.x);
    if (T >= g.ntiles) return;
    const uint32_t t0 = sub[(size_t)T * NS], t1 = sub[(size_t)(T + 1) * NS];
    if (t1 == t0) {
        // an empty tile: zero ghost partials (the fold reads every tile's;
        // a per-tile flag that let the fold skip them cost it more, 35 -> 45
        // us on the dense box, than the zero stores cost here)
        Real4<Real>* out = reinterpret_cast<Real4<Real>*>(partial) + (size_t)T * g.next;
        Real4<Real> z;
        z.v[0] = z.v[1] = z.v[2] = z.v[3] = Real(0);
        for (int e = threadIdx.x; e < g.next; e += blockDim.x) out[e] = z;
        return;
    }
    if (t1 - t0 >= dense_min) {
        deposit_dense_tile(g, A, sub, glist, gcount, partial, T, S, R);
        return;
    }
    const int ls = threadIdx.x;
    const uint32_t key = (uint32_t)(T * NS + ls);
    const uint32_t p0 = sub[key], p1 = sub[key + 1];
    const int o = ls & 7, cell = ls >> 3;
    const bool hx = o & 1, hy = (o >> 1) & 1, hz = (o >> 2) & 1;
    Real acc[4][8];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[q][c] = Real(0);
    // The tile's particles go through LDS in chunks read with coalesced
    // loads; each thread then takes its subcell's particles of the chunk, in
    // particle order (the same sums as reading them directly).  A particle is
    // staged as (x, y, z, vx) and (vy, vz): two LDS reads instead of six.
    const int CH = deposit_chunk<Real>(g);
    Real4<Real>* L4 = reinterpret_cast<Real4<Real>*>(S);          // x, y, z, vx
    Real2<Real>* L2 = reinterpret_cast<Real2<Real>*>(S + 4 * CH);  // vy, vz
    for (uint32_t cs = t0; cs < t1; cs += (uint32_t)CH) {
        const int n = (int)min((uint32_t)CH, t1 - cs);
        __syncthreads();  // the previous chunk is consumed
        for (int q = ls; q < n; q += blockDim.x) {
            const uint32_t p = cs + (uint32_t)q;
            Real4<Real> r;
            r.v[0] = A.x[p]; r.v[1] = A.y[p]; r.v[2] = A.z[p]; r.v[3] = A.vx[p];
            Real2<Real> t;
            t.v[0] = A.vy[p]; t.v[1] = A.vz[p];
            L4[q] = r;
            L2[q] = t;
        }
        __syncthreads();
        const uint32_t a = max(p0, cs), b = min(p1, cs + (uint32_t)n);
        for (uint32_t p = a; p < b; ++p) {
            const int q = (int)(p - cs);
            const Real4<Real> r = L4[q];
            const Real2<Real> t = L2[q];
            cic_accumulate(acc, hx, hy, hz, r.v[0], r.v[1], r.v[2], r.v[3], t.v[0], t.v[1]);
        }
    }
    // the eight octant lanes of each cell: 64 corner sums -> 27 targets
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int r = 0; r < 4; ++r) octant_stage<1>(acc[q][2 * r], acc[q][2 * r + 1], hx);
#pragma unroll
        for (int r = 0; r < 4; ++r) octant_stage<2>(acc[q][(r & 1) + 4 * (r >> 1)], acc[q][(r & 1) + 4 * (r >> 1) + 2], hy);
#pragma unroll
        for (int r = 0; r < 4; ++r) octant_stage<4>(acc[q][r], acc[q][r + 4], hz);
    }
    __syncthreads();  // the staging is read; the target planes may be written
    const int CB = OB, QP = 27 * CB;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int sx = c & 1, sy = (c >> 1) & 1, sz = c >> 2;
        if ((hx && sx) || (hy && sy) || (hz && sz)) continue;  // slot unused by this octant
        const int dx = hx ? 1 : sx - 1, dy = hy ? 1 : sy - 1, dz = hz ? 1 : sz - 1;
        const int t27 = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
#pragma unroll
        for (int q = 0; q < 4; ++q) S[q * QP + t27 * CB + cell] = acc[q][c];
    }
    if (ls < 4) S[ls * QP + g.tc] = Real(0);  // the lists' padding entries read this word
    __syncthreads();
    Real4<Real>* out = reinterpret_cast<Real4<Real>*>(partial) + (size_t)T * g.next;
    for (int e = ls; e < g.next; e += blockDim.x) {
        // the cell terms reaching this cell: a list shared by all tiles, read
        // eight entries at a time (the padding adds zeros after the terms)
        const int ng = (tcount[e] + 7) >> 3;
        const uint4* lst = reinterpret_cast<const uint4*>(tlist + (size_t)e * 32);
        Real4<Real> r;
        r.v[0] = r.v[1] = r.v[2] = r.v[3] = Real(0);
        for (int gi = 0; gi < ng; ++gi) {
            const uint4 l8 = lst[gi];
            const uint32_t w[4] = {l8.x, l8.y, l8.z, l8.w};
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int off = (int)((w[t >> 1] >> (16 * (t & 1))) & 0xFFFFu);
                r.v[0] += S[off];
                r.v[1] += S[QP + off];
                r.v[2] += S[2 * QP + off];
                r.v[3] += S[3 * QP + off];
            }
        }
        out[e] = r;
    }
}

// Accumulator::addRemoteUpdates (src/particle_accumulator.cpp:260-335): each
// cell sums in FP64, species by species and in a fixed tile order, the ghost
// partials of every tile whose extended region holds it, times the species'
// w*q; then rho/=dV with the rho floor, J/=dV.
constexpr int MAX_FOLD_SPECIES = 8;
template <typename Real>
struct FoldSources {
    const Real* partial[MAX_FOLD_SPECIES];
    double wq[MAX_FOLD_SPECIES];
    int n;
};

template <typename Real>
__global__ __launch_bounds__(256) void k_fold(Geom g, FoldSources<Real> src, double inv_dV, double min_rho,
                                             Real* __restrict__ rho, Real* __restrict__ jx, Real* __restrict__ jy,
                                             Real* __restrict__ jz, unsigned long long* __restrict__ counters) {
    const int i = blockIdx.x * 32 + threadIdx.x, j = blockIdx.y * 8 + threadIdx.y, k = blockIdx.z;
    bool floored = false;
    if (i < g.nx && j < g.ny) {
        const int n = cell_index(g, i, j, k);
        int tx[6], ex[6], ty[6], ey[6], tz[6], ez[6];
        const bool fx = !g.per[0] || g.ntx >= 3, fy = !g.per[1] || g.nty >= 3, fz = !g.per[2] || g.ntz >= 3;
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        if (fx && fy && fz && g.tx >= 2 && g.ty >= 2 && g.tz >= 2) {
            // at most two candidate tiles per axis: all eight loads of a
            // species issued before its sum (the same order as below)
            const int mx = axis_candidates_fast(i, g.ltx, g.ntx, g.per[0], tx, ex);
            const int my = axis_candidates_fast(j, g.lty, g.nty, g.per[1], ty, ey);
            const int mz = axis_candidates_fast(k, g.ltz, g.ntz, g.per[2], tz, ez);
            for (int s = 0; s < src.n; ++s) {
                const Real4<Real>* part = reinterpret_cast<const Real4<Real>*>(src.partial[s]);
                Real4<Real> v[8];
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
#pragma unroll
                        for (int a = 0; a < 2; ++a) {
                            const int q = a + 2 * b + 4 * c;
                            if (a < mx && b < my && c < mz) {
                                const int TT = tx[a] + g.ntx * (ty[b] + g.nty * tz[c]);
                                const int e = ex[a] + g.ex * (ey[b] + g.ey * ez[c]);
                                v[q] = part[(size_t)TT * g.next + e];
                            }
                        }
                double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
#pragma unroll
                        for (int a = 0; a < 2; ++a) {
                            const int q = a + 2 * b + 4 * c;
                            if (!(a < mx && b < my && c < mz)) continue;
                            t0 += (double)v[q].v[0]; t1 += (double)v[q].v[1];
                            t2 += (double)v[q].v[2]; t3 += (double)v[q].v[3];
                        }
                s0 += src.wq[s] * t0; s1 += src.wq[s] * t1; s2 += src.wq[s] * t2; s3 += src.wq[s] * t3;
            }
        } else {
            const int mx = axis_candidates(i, g.nx, g.ltx, g.ntx, g.per[0], tx, ex);
            const int my = axis_candidates(j, g.ny, g.lty, g.nty, g.per[1], ty, ey);
            const int mz = axis_candidates(k, g.nz, g.ltz, g.ntz, g.per[2], tz, ez);
            for (int s = 0; s < src.n; ++s) {
                double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
                for (int c = 0; c < mz; ++c)
                    for (int b = 0; b < my; ++b)
                        for (int a = 0; a < mx; ++a) {
                            const int TT = tx[a] + g.ntx * (ty[b] + g.nty * tz[c]);
                            const int e = ex[a] + g.ex * (ey[b] + g.ey * ez[c]);
                            const Real4<Real> v = reinterpret_cast<const Real4<Real>*>(src.partial[s])[(size_t)TT * g.next + e];
                            t0 += (double)v.v[0]; t1 += (double)v.v[1]; t2 += (double)v.v[2]; t3 += (double)v.v[3];
                        }
                s0 += src.wq[s] * t0; s1 += src.wq[s] * t1; s2 += src.wq[s] * t2; s3 += src.wq[s] * t3;
            }
        }
        double r = s0 * inv_dV;
        if (r < min_rho) { r = min_rho; floored = true; }
        rho[n] = (Real)r;
        jx[n] = (Real)(s1 * inv_dV);
        jy[n] = (Real)(s2 * inv_dV);
        jz[n] = (Real)(s3 * inv_dV);
    }
    count_event(counters, C_MIN_CELL_RHO, floored);
}

// WRITE_GRID_TEMPORAL_AVERAGES of the accumulator (src/particle_accumulator.
// cpp:148-151, 287-300): one species' charge and current sums of this step
// (its tile partials folded in k_fold's fixed order, times w q, before /dV
// and the floor) added to its running FP64 sums nave[n], vave[l*nc + n].
template <typename Real>
__global__ __launch_bounds__(256) void k_average_moments(Geom g, const Real* __restrict__ partial, double wq,
                                                        double* __restrict__ nave, double* __restrict__ vave) {
    const int i = blockIdx.x * 32 + threadIdx.x, j = blockIdx.y * 8 + threadIdx.y, k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int n = cell_index(g, i, j, k);
    int tx[6], ex[6], ty[6], ey[6], tz[6], ez[6];
    const int mx = axis_candidates(i, g.nx, g.ltx, g.ntx, g.per[0], tx, ex);
    const int my = axis_candidates(j, g.ny, g.lty, g.nty, g.per[1], ty, ey);
    const int mz = axis_candidates(k, g.nz, g.ltz, g.ntz, g.per[2], tz, ez);
    double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
    for (int c = 0; c < mz; ++c)
        for (int b = 0; b < my; ++b)
            for (int a = 0; a < mx; ++a) {
                const int TT = tx[a] + g.ntx * (ty[b] + g.nty * tz[c]);
                const int e = ex[a] + g.ex * (ey[b] + g.ey * ez[c]);
                const Real4<Real> v = reinterpret_cast<const Real4<Real>*>(partial)[(size_t)TT * g.next + e];
                t0 += (double)v.v[0]; t1 += (double)v.v[1]; t2 += (double)v.v[2]; t3 += (double)v.v[3];
            }
    nave[n] += wq * t0;
    vave[n] += wq * t1;
    vave[(size_t)g.nc + n] += wq * t2;
    vave[2 * (size_t)g.nc + n] += wq * t3;
}

// cellAverageB += cellB (src/hybrid_propagator.cpp:157-167), FP64 sums
template <typename Real>
__global__ __launch_bounds__(256) void k_average_cellB(int nc, const Real* __restrict__ b, double* __restrict__ ave) {
    const size_t q = blockIdx.x * (size_t)256 + threadIdx.x;
    if (q < 3 * (size_t)nc) ave[q] += (double)b[q];
}

// ParticleBoundaryCondHybrid::apply (src/particle_boundary_cond_hybrid.h:
// 112-185): empty exterior cells, remove particles inside the obstacle in
// innerFlagParticle cells.  Removed particles get x = -1; the next sort drops
// them.  16 lanes per listed cell (four cells a wave), striding through the
// cell's particles.
constexpr int BC_LANES = 16;
constexpr int BC_CELLS_PER_BLOCK = 256 / BC_LANES;
// The removed particles are counted for the logs (logCounterParticleEscape
// / Impact and their kinetic energies, :121-122,153-154): per listed cell,
// acc[4q..4q+3] += (escapes, sum v^2 of escapes, impacts, sum v^2 of
// impacts), summed over the cell's lanes in a fixed butterfly.
template <typename Real>
__global__ __launch_bounds__(256) void k_boundary(Geom g, const int* __restrict__ cells, int ncells,
                                                 const uint32_t* __restrict__ sub, PArrays<Real> A,
                                                 double R2, double* __restrict__ acc) {
    const int q = blockIdx.x * BC_CELLS_PER_BLOCK + (int)(threadIdx.x / BC_LANES);
    const int lane = threadIdx.x & (BC_LANES - 1);
    const bool valid = q < ncells;
    double c[4] = {0, 0, 0, 0};
    if (valid) {
        const int n = cells[q];
        int i, j, k;
        cell_ijk(g, n, i, j, k);
        const uint32_t kc = cell_key(g, i, j, k);
        const uint32_t p0 = sub[(size_t)kc * 8], p1 = sub[(size_t)kc * 8 + 8];
        const bool exterior = !all_nbrs(g, i, j, k);
        const double xb = g.x0 + i * g.dx, yb = g.y0 + j * g.dx, zb = g.z0 + k * g.dx;
        for (uint32_t p = p0 + lane; p < p1; p += BC_LANES) {
            if (A.x[p] < Real(0)) continue;
            const double vx = A.vx[p], vy = A.vy[p], vz = A.vz[p];
            const double v2 = vx * vx + vy * vy + vz * vz;
            if (exterior) {
                A.x[p] = Real(-1);
                c[0] += 1.0; c[1] += v2;
                continue;
            }
            const double x = xb + (double)A.x[p] * g.dx, y = yb + (double)A.y[p] * g.dx, z = zb + (double)A.z[p] * g.dx;
            if (x * x + y * y + z * z < R2) {
                A.x[p] = Real(-1);
                c[2] += 1.0; c[3] += v2;
            }
        }
    }
    // most listed cells remove nothing in a step: then the wave is done
    if (!__ballot(c[0] != 0.0 || c[2] != 0.0)) return;
    for (int t = 0; t < 4; ++t)
        for (int off = BC_LANES / 2; off > 0; off >>= 1) c[t] += __shfl_xor(c[t], off, 64);
    if (valid && lane == 0)
        for (int t = 0; t < 4; ++t) acc[(size_t)q * 4 + t] += c[t];
}

// FP64 particle sums (logCalcParticle, src/diagnostics.h:209-228), one
// partial per block, summed on the host in block order: count, sum vx, vy,
// vz, sum |v|, sum v^2 (unweighted; one weight per species) and max |v|.
constexpr int DIAG_BLOCKS = 1024;
constexpr int DIAG_Q = 7;
// Injected particles waiting past the store's end (*extra of them, extra may
// be null) are included.
template <typename Real>
__global__ __launch_bounds__(256) void k_diag(PArrays<Real> A, const uint32_t* __restrict__ np_ptr,
                                             const uint32_t* __restrict__ extra, uint32_t cap,
                                             double* __restrict__ partial) {
    __shared__ double red[DIAG_Q][256];
    uint32_t np = *np_ptr;
    if (extra) np = (uint64_t)np + *extra > cap ? cap : np + *extra;
    double s[DIAG_Q] = {0, 0, 0, 0, 0, 0, 0};
    for (uint32_t p = blockIdx.x * 256u + threadIdx.x; p < np; p += 256u * gridDim.x) {
        if (A.x[p] < Real(0)) continue;
        const double vx = A.vx[p], vy = A.vy[p], vz = A.vz[p];
        const double v2 = vx * vx + vy * vy + vz * vz, v = sqrt(v2);
        s[0] += 1.0; s[1] += vx; s[2] += vy; s[3] += vz; s[4] += v; s[5] += v2;
        s[6] = fmax(s[6], v);
    }
    for (int q = 0; q < DIAG_Q; ++q) red[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            for (int q = 0; q < DIAG_Q - 1; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + off];
            red[DIAG_Q - 1][threadIdx.x] = fmax(red[DIAG_Q - 1][threadIdx.x], red[DIAG_Q - 1][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < DIAG_Q; ++q) partial[blockIdx.x * DIAG_Q + q] = red[q][0];
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, T v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)blockDim.x * gridDim.x) p[i] = v;
}

}  // namespace rh
