// rh_particles.hpp -- device kernels of the particle half of the step:
// Boris-Buneman push, the cell migration as a deterministic counting sort by
// subcell, the cloud-in-cell deposit with fixed-order reductions, the FP64
// cross-tile ghost fold, particle boundary conditions and FP64 diagnostics.
//
// Storage (per species, structure of arrays, Real = float or double):
//   x,y,z    cell-local position in units of dx, in [0,1)
//   vx,vy,vz velocity in m/s
//   cell     global cell index n, DEAD_CELL once removed
// sorted by subcell key (rh_geom.hpp); sub[k] .. sub[k+1] is the range of
// subcell key k.
#pragma once

#include <hip/hip_runtime.h>

#include "rh_geom.hpp"

namespace rh {

template <typename Real>
struct PArrays {
    Real *x, *y, *z, *vx, *vy, *vz;
    uint32_t* cell;
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
};

enum DevCounter { C_MAX_VI = 0, C_MIN_CELL_RHO, C_MAX_CELL_UE, C_MIN_NODE_RHO, C_MAX_NODE_UE,
                  C_MAX_E, C_MAX_VW, C_DROPPED, C_RANK_OVERFLOW, C_N };

__device__ __forceinline__ void count_event(unsigned long long* counters, int id, bool pred) {
    // one atomic per wave: integer, so the total does not depend on order
    const unsigned long long b = __ballot(pred);
    if (b && (threadIdx.x & 63) == (unsigned)__builtin_ctzll(b))
        atomicAdd(counters + id, (unsigned long long)__popcll(b));
}

template <typename Real> __device__ __forceinline__ Real below_one();
template <> __device__ __forceinline__ float below_one<float>() { return 0x1.fffffep-1f; }
template <> __device__ __forceinline__ double below_one<double>() { return 0x1.fffffffffffffp-1; }

// One workgroup (one wave) per tile.  Pushes the tile's particles
// (BorisBuneman::propagate, src/particle_propagator_boris_buneman.h:64-194,
// with getFields/face2r/cell2r, src/hybrid_propagator.cpp:1740-1891), moves
// them in FP64 (the cell crossing), and ranks each particle among the
// particles of the tile that go to the same subcell slot of the tile's
// extended region, in memory order.  Writes the pushed state to B at the same
// index, packed (slot << 20 | rank), and the tile's slot counts.
template <typename Real>
__global__ __launch_bounds__(64) void k_push_rank(Geom g, PushParams pp, PushFields<Real> f,
                                                 PArrays<Real> A, PArrays<Real> B,
                                                 uint32_t* __restrict__ packed,
                                                 const uint32_t* __restrict__ sub,
                                                 uint32_t* __restrict__ cnt_tile,
                                                 unsigned long long* __restrict__ counters, int kbits) {
    extern __shared__ uint32_t cnt[];
    const int T = blockIdx.x;
    const int lane = threadIdx.x;
    for (int e = lane; e < g.extk; e += 64) cnt[e] = 0;
    __syncthreads();
    int ox, oy, oz;
    tile_origin(g, T, ox, oy, oz);
    const uint32_t p0 = sub[(size_t)T * g.tc * 8];
    const uint32_t p1 = sub[(size_t)(T + 1) * g.tc * 8];
    const uint32_t deadkey = (uint32_t)MAX_EXTK;
    for (uint32_t base = p0; base < p1; base += 64) {
        const uint32_t p = base + lane;
        const bool act = p < p1;
        uint32_t key = deadkey;
        bool clamped = false, dropped = false;
        if (act) {
            const uint32_t n = A.cell[p];
            Real X = A.x[p], Y = A.y[p], Z = A.z[p];
            Real vx = A.vx[p], vy = A.vy[p], vz = A.vz[p];
            if (n != DEAD_CELL) {
                int i, j, k;
                cell_ijk(g, (int)n, i, j, k);
                bool accel = pp.accelerate && all_nbrs(g, i, j, k) && !pp.initial_flow_through;
                if (accel) {
                    const int nmx = nbr(g, i, j, k, -1, 0, 0);
                    const int nmy = nbr(g, i, j, k, 0, -1, 0);
                    const int nmz = nbr(g, i, j, k, 0, 0, -1);
                    // face2r: each component linear along its own axis only
                    const Real Bx = X * f.fbx[n] + (Real(1) - X) * f.fbx[nmx];
                    const Real By = Y * f.fby[n] + (Real(1) - Y) * f.fby[nmy];
                    const Real Bz = Z * f.fbz[n] + (Real(1) - Z) * f.fbz[nmz];
                    // cell2r: nearest grid point
                    const Real Ux = f.uex[n], Uy = f.uey[n], Uz = f.uez[n];
                    Real Ex = By * Uz - Bz * Uy;
                    Real Ey = Bz * Ux - Bx * Uz;
                    Real Ez = Bx * Uy - By * Ux;
                    if (pp.use_e_pressure) { Ex += f.epx[n]; Ey += f.epy[n]; Ez += f.epz[n]; }
                    const Real qm = (Real)pp.qmdt2;
                    const Real dvx = qm * Ex, dvy = qm * Ey, dvz = qm * Ez;
                    const Real tx = qm * Bx, ty = qm * By, tz = qm * Bz;
                    const Real t2 = tx * tx + ty * ty + tz * tz;
                    const Real b2 = Real(2) / (Real(1) + t2);
                    const Real sx = b2 * tx, sy = b2 * ty, sz = b2 * tz;
                    const Real vmx = vx + dvx, vmy = vy + dvy, vmz = vz + dvz;
                    const Real v0x = vmx + vmy * tz - vmz * ty;
                    const Real v0y = vmy + vmz * tx - vmx * tz;
                    const Real v0z = vmz + vmx * ty - vmy * tx;
                    vx = vmx + v0y * sz - v0z * sy + dvx;
                    vy = vmy + v0z * sx - v0x * sz + dvy;
                    vz = vmz + v0x * sy - v0y * sx + dvz;
                    if (pp.use_gravity) {
                        const double rx = g.x0 + (i + (double)X) * g.dx;
                        const double ry = g.y0 + (j + (double)Y) * g.dx;
                        const double rz = g.z0 + (k + (double)Z) * g.dx;
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
                // move in FP64; the integer part is the cell crossing
                double Xd = (double)X + (double)vx * pp.dt_dx;
                double Yd = (double)Y + (double)vy * pp.dt_dx;
                double Zd = (double)Z + (double)vz * pp.dt_dx;
                const double fx = floor(Xd), fy = floor(Yd), fz = floor(Zd);
                int sx = (int)fx, sy = (int)fy, sz = (int)fz;
                bool ok = (fx >= -1.0 && fx <= 1.0 && fy >= -1.0 && fy <= 1.0 && fz >= -1.0 && fz <= 1.0);
                int nn = -1;
                if (ok) nn = nbr(g, i, j, k, sx, sy, sz);
                if (nn >= 0) {
                    X = (Real)(Xd - fx); Y = (Real)(Yd - fy); Z = (Real)(Zd - fz);
                    if (X >= Real(1)) X = below_one<Real>();
                    if (Y >= Real(1)) Y = below_one<Real>();
                    if (Z >= Real(1)) Z = below_one<Real>();
                    const int o = (X >= Real(0.5)) + 2 * (Y >= Real(0.5)) + 4 * (Z >= Real(0.5));
                    const int ex = i - ox + sx + 1, ey = j - oy + sy + 1, ez = k - oz + sz + 1;
                    key = (uint32_t)((ex + g.ex * (ey + g.ey * ez)) * 8 + o);
                } else {
                    dropped = true;
                }
            }
            B.x[p] = X; B.y[p] = Y; B.z[p] = Z;
            B.vx[p] = vx; B.vy[p] = vy; B.vz[p] = vz;
        }
        // stable rank among the lanes holding the same slot
        unsigned long long match = __ballot(act && key != deadkey);
        for (int b = 0; b < kbits; ++b) {
            const bool bit = (key >> b) & 1u;
            const unsigned long long bal = __ballot(bit);
            match &= bit ? bal : ~bal;
        }
        const unsigned long long below = (lane == 0) ? 0ull : (match & ((~0ull) >> (64 - lane)));
        const uint32_t rin = (uint32_t)__popcll(below);
        const uint32_t gsz = (uint32_t)__popcll(match);
        if (act) {
            if (key != deadkey) {
                const uint32_t c0 = cnt[key];
                const uint32_t rank = c0 + rin;
                if (rin == gsz - 1) cnt[key] = c0 + gsz;
                packed[p] = (rank <= RANK_MASK) ? ((key << RANK_BITS) | rank) : DEAD_PACKED;
                if (rank > RANK_MASK) atomicAdd(counters + C_RANK_OVERFLOW, 1ull);
            } else {
                packed[p] = DEAD_PACKED;
            }
        }
        count_event(counters, C_MAX_VI, clamped);
        count_event(counters, C_DROPPED, dropped);
    }
    __syncthreads();
    for (int e = lane; e < g.extk; e += 64) cnt_tile[(size_t)T * g.extk + e] = cnt[e];
}

// decode a subcell key into tile, cell coordinates and octant
__device__ __forceinline__ void decode_key(const Geom& g, uint32_t key, int& T, int& i, int& j, int& k, int& o) {
    T = (int)(key / (uint32_t)(g.tc * 8));
    const int r = (int)(key % (uint32_t)(g.tc * 8));
    o = r & 7;
    const int lc = r >> 3;
    int ox, oy, oz;
    tile_origin(g, T, ox, oy, oz);
    i = ox + lc % g.tx;
    j = oy + (lc / g.tx) % g.ty;
    k = oz + lc / (g.tx * g.ty);
}

// Number of particles arriving in subcell key k: the slot counts of every
// tile whose extended region holds the cell, in a fixed order.  With
// write_base, the slot counts are replaced by the first destination index of
// each slot (the scan of totals must already be in sub).
template <bool WRITE_BASE>
__global__ __launch_bounds__(256) void k_slot_offsets(Geom g, uint32_t* __restrict__ cnt_tile,
                                                     uint32_t* __restrict__ out) {
    const uint32_t key = blockIdx.x * 256u + threadIdx.x;
    if (key >= g.nkeys) return;
    int T, i, j, k, o;
    decode_key(g, key, T, i, j, k, o);
    int tx[6], ex[6], ty[6], ey[6], tz[6], ez[6];
    const int mx = axis_candidates(i, g.nx, g.tx, g.ntx, g.per[0], tx, ex);
    const int my = axis_candidates(j, g.ny, g.ty, g.nty, g.per[1], ty, ey);
    const int mz = axis_candidates(k, g.nz, g.tz, g.ntz, g.per[2], tz, ez);
    uint32_t run = WRITE_BASE ? out[key] : 0u;
    for (int c = 0; c < mz; ++c)
        for (int b = 0; b < my; ++b)
            for (int a = 0; a < mx; ++a) {
                const int TT = tx[a] + g.ntx * (ty[b] + g.nty * tz[c]);
                const size_t slot = (size_t)TT * g.extk + (size_t)((ex[a] + g.ex * (ey[b] + g.ey * ez[c])) * 8 + o);
                const uint32_t v = cnt_tile[slot];
                if (WRITE_BASE) cnt_tile[slot] = run;
                run += v;
            }
    if (!WRITE_BASE) out[key] = run;
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
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0;
    for (int q = 0; q < SCAN_ITEMS; ++q)
        if (base + q < n) s += in[base + q];
    uint32_t tot;
    block_exclusive_scan(s, smem, tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const uint32_t* __restrict__ in, uint32_t n,
                                                            const uint32_t* __restrict__ boff,
                                                            uint32_t* __restrict__ out) {
    __shared__ uint32_t smem[SCAN_THREADS];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        v[q] = (base + q < n) ? in[base + q] : 0u;
        s += v[q];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, smem, tot) + boff[blockIdx.x];
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        if (base + q < n) out[base + q] = run;
        run += v[q];
    }
    // the grand total goes one past the end
    if (base + SCAN_ITEMS >= n && base < n) out[n] = run;
}

// Moves every particle of an old tile to its slot: base of the slot (written
// by k_slot_offsets<true>) plus its rank.  One wave per old tile.
template <typename Real>
__global__ __launch_bounds__(64) void k_scatter(Geom g, PArrays<Real> B, const uint32_t* __restrict__ packed,
                                               const uint32_t* __restrict__ cnt_tile,
                                               const uint32_t* __restrict__ sub_old, PArrays<Real> A,
                                               uint32_t cap, unsigned long long* __restrict__ counters) {
    const int T = blockIdx.x;
    const int lane = threadIdx.x;
    int ox, oy, oz;
    tile_origin(g, T, ox, oy, oz);
    const uint32_t p0 = sub_old[(size_t)T * g.tc * 8];
    const uint32_t p1 = sub_old[(size_t)(T + 1) * g.tc * 8];
    for (uint32_t p = p0 + lane; p < p1; p += 64) {
        const uint32_t pk = packed[p];
        if (pk == DEAD_PACKED) continue;
        const uint32_t slot = pk >> RANK_BITS;
        const uint32_t dst = cnt_tile[(size_t)T * g.extk + slot] + (pk & RANK_MASK);
        if (dst >= cap) {  // cannot happen with consistent counts; never write out of bounds
            atomicAdd(counters + C_RANK_OVERFLOW, 1ull);
            continue;
        }
        const int e = (int)(slot >> 3);
        const int ex = e % g.ex, ey = (e / g.ex) % g.ey, ez = e / (g.ex * g.ey);
        const int i = wrap_axis(ox + ex - 1, g.nx, g.per[0]);
        const int j = wrap_axis(oy + ey - 1, g.ny, g.per[1]);
        const int k = wrap_axis(oz + ez - 1, g.nz, g.per[2]);
        A.x[dst] = B.x[p]; A.y[dst] = B.y[p]; A.z[dst] = B.z[p];
        A.vx[dst] = B.vx[p]; A.vy[dst] = B.vy[p]; A.vz[dst] = B.vz[p];
        A.cell[dst] = (uint32_t)cell_index(g, i, j, k);
    }
}

// Cloud-in-cell deposit (Accumulator::accumulateCell,
// src/particle_accumulator.cpp:43-156).  One workgroup per tile, one thread
// per subcell: a subcell's particles share one 2x2x2 footprint, so the thread
// sums, in particle order, the 8 corner weights times (1, vx, vy, vz).  The
// subcell sums are staged in LDS and every cell of the tile's extended region
// adds the (up to 64) subcell terms that reach it in a fixed order.  The tile
// result, times w*q, is added in FP64 to the tile's own ghost buffer; no
// atomics, so the result is the same bit for bit on every run.
template <typename Real>
__global__ __launch_bounds__(512) void k_deposit(Geom g, PArrays<Real> A, const uint32_t* __restrict__ sub,
                                                double wq, double* __restrict__ tile_acc) {
    extern __shared__ unsigned char smem_raw[];
    Real* S = reinterpret_cast<Real*>(smem_raw);
    const int NS = g.tc * 8;
    const int T = blockIdx.x;
    const int ls = threadIdx.x;
    {
        const uint32_t key = (uint32_t)(T * NS + ls);
        const uint32_t p0 = sub[key], p1 = sub[key + 1];
        const int o = ls & 7;
        Real acc[4][8];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < 8; ++c) acc[q][c] = Real(0);
        const bool hx = o & 1, hy = (o >> 1) & 1, hz = (o >> 2) & 1;
        for (uint32_t p = p0; p < p1; ++p) {
            const Real X = A.x[p], Y = A.y[p], Z = A.z[p];
            const Real vx = A.vx[p], vy = A.vy[p], vz = A.vz[p];
            // weights of the lower / upper target along each axis
            const Real wx1 = hx ? X - Real(0.5) : X + Real(0.5);
            const Real wy1 = hy ? Y - Real(0.5) : Y + Real(0.5);
            const Real wz1 = hz ? Z - Real(0.5) : Z + Real(0.5);
            const Real wx0 = Real(1) - wx1, wy0 = Real(1) - wy1, wz0 = Real(1) - wz1;
            const Real wx[2] = {wx0, wx1}, wy[2] = {wy0, wy1}, wz[2] = {wz0, wz1};
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const Real w = wx[c & 1] * wy[(c >> 1) & 1] * wz[c >> 2];
                acc[0][c] += w;
                acc[1][c] += w * vx;
                acc[2][c] += w * vy;
                acc[3][c] += w * vz;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < 8; ++c) S[(q * 8 + c) * NS + ls] = acc[q][c];
    }
    __syncthreads();
    const int nout = g.next * 4;
    double* out = tile_acc + (size_t)T * nout;
    for (int idx = ls; idx < nout; idx += blockDim.x) {
        const int e = idx >> 2, q = idx & 3;
        const int lx = e % g.ex - 1, ly = (e / g.ex) % g.ey - 1, lz = e / (g.ex * g.ey) - 1;
        Real sum = Real(0);
        // target local coordinate l = cell + half - 1 + corner, half/corner in {0,1}
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
                                const int lsub = (cx + g.tx * (cy + g.ty * cz)) * 8 + hx + 2 * hy + 4 * hz;
                                const int corner = kx + 2 * ky + 4 * kz;
                                sum += S[(q * 8 + corner) * NS + lsub];
                            }
                    }
            }
        out[idx] += wq * (double)sum;
    }
}

// Accumulator::addRemoteUpdates (src/particle_accumulator.cpp:260-335): each
// cell sums, in FP64 and a fixed order, the ghost contributions of every tile
// whose extended region holds it, then rho/=dV with the rho floor, J/=dV.
template <typename Real>
__global__ __launch_bounds__(256) void k_fold(Geom g, const double* __restrict__ tile_acc, double inv_dV,
                                             double min_rho, Real* __restrict__ rho, Real* __restrict__ jx,
                                             Real* __restrict__ jy, Real* __restrict__ jz,
                                             unsigned long long* __restrict__ counters) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    bool floored = false;
    if (n < g.nc) {
        int i, j, k;
        cell_ijk(g, n, i, j, k);
        int tx[6], ex[6], ty[6], ey[6], tz[6], ez[6];
        const int mx = axis_candidates(i, g.nx, g.tx, g.ntx, g.per[0], tx, ex);
        const int my = axis_candidates(j, g.ny, g.ty, g.nty, g.per[1], ty, ey);
        const int mz = axis_candidates(k, g.nz, g.tz, g.ntz, g.per[2], tz, ez);
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        for (int c = 0; c < mz; ++c)
            for (int b = 0; b < my; ++b)
                for (int a = 0; a < mx; ++a) {
                    const int TT = tx[a] + g.ntx * (ty[b] + g.nty * tz[c]);
                    const int e = ex[a] + g.ex * (ey[b] + g.ey * ez[c]);
                    const double* v = tile_acc + ((size_t)TT * g.next + e) * 4;
                    s0 += v[0]; s1 += v[1]; s2 += v[2]; s3 += v[3];
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

// ParticleBoundaryCondHybrid::apply (src/particle_boundary_cond_hybrid.h:
// 112-185): empty exterior cells, remove particles inside the obstacle in
// innerFlagParticle cells.  Removed particles get cell = DEAD_CELL; the next
// sort drops them.  One thread per listed cell.
template <typename Real>
__global__ __launch_bounds__(256) void k_boundary(Geom g, const int* __restrict__ cells, int ncells,
                                                 const uint32_t* __restrict__ sub, PArrays<Real> A,
                                                 double R2) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= ncells) return;
    const int n = cells[q];
    int i, j, k;
    cell_ijk(g, n, i, j, k);
    const uint32_t kc = cell_key(g, i, j, k);
    const uint32_t p0 = sub[(size_t)kc * 8], p1 = sub[(size_t)kc * 8 + 8];
    const bool exterior = !all_nbrs(g, i, j, k);
    const double xb = g.x0 + i * g.dx, yb = g.y0 + j * g.dx, zb = g.z0 + k * g.dx;
    for (uint32_t p = p0; p < p1; ++p) {
        if (exterior) { A.cell[p] = DEAD_CELL; continue; }
        const double x = xb + (double)A.x[p] * g.dx, y = yb + (double)A.y[p] * g.dx, z = zb + (double)A.z[p] * g.dx;
        if (x * x + y * y + z * z < R2) A.cell[p] = DEAD_CELL;
    }
}

// FP64 conservation sums (logCalcParticle, src/diagnostics.h:209-220), one
// partial per block, summed on the host in block order.
constexpr int DIAG_BLOCKS = 1024;
template <typename Real>
__global__ __launch_bounds__(256) void k_diag(PArrays<Real> A, const uint32_t* __restrict__ np_ptr,
                                             double* __restrict__ partial) {
    __shared__ double red[5][256];
    const uint32_t np = *np_ptr;
    double s[5] = {0, 0, 0, 0, 0};
    for (uint32_t p = blockIdx.x * 256u + threadIdx.x; p < np; p += 256u * gridDim.x) {
        if (A.cell[p] == DEAD_CELL) continue;
        const double vx = A.vx[p], vy = A.vy[p], vz = A.vz[p];
        s[0] += 1.0; s[1] += vx; s[2] += vy; s[3] += vz; s[4] += vx * vx + vy * vy + vz * vz;
    }
    for (int q = 0; q < 5; ++q) red[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int q = 0; q < 5; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 5; ++q) partial[blockIdx.x * 5 + q] = red[q][0];
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, T v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)blockDim.x * gridDim.x) p[i] = v;
}

}  // namespace rh
