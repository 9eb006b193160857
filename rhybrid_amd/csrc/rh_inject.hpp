// rh_inject.hpp -- particle injection on the device (src/particle_injector.cpp).
//
// Random numbers come from Philox4x32-10 (Salmon, Moraes, Dror and Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11), a counter-based
// generator: key = the injector's seed, counter = (cell, timestep, particle,
// block).  Every cell and particle has its own stream, so the particles of a
// step do not depend on the order threads run in, and the FP64 oracle
// reproduces them.  The reference draws from Corsair's sequential generator
// instead; the samplers on top follow the reference:
//   probRound            src/particle_injector.cpp:41-47
//   sampleGaussian       :59-80 (polar method, second deviate kept per stream)
//   sampleDerivGaussian  :97-114 (rejection from the unit Gaussian at the mode)
//
// New particles are written past the end of the sorted store in the order of
// the listed cells.  The list is sorted by subcell key, so each tile's new
// particles are one range; the next push/sort pass of the tile takes them
// after the tile's own particles.  The solar wind gives every listed cell a
// fixed number of places (at most ceil(rate) particles per step), filled in
// one launch; the exosphere's per-cell rates differ by orders of magnitude,
// so its step is count (new particles per listed cell), scan (offsets), fill.
#pragma once

#include <hip/hip_runtime.h>

#include "rh_geom.hpp"
#include "rh_particles.hpp"

namespace rh {

RH_HD void philox4x32_10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
    }
}

// 53 random bits -> [0,1)
RH_HD double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

struct RngStream {
    uint32_t k0, k1, c0, c1, c2, blk;
    double su, sg;
    bool hu, hg;
    RH_HD RngStream(uint32_t key0, uint32_t key1, uint32_t a, uint32_t b, uint32_t c)
        : k0(key0), k1(key1), c0(a), c1(b), c2(c), blk(0), su(0), sg(0), hu(false), hg(false) {}
    RH_HD double uniform() {
        if (hu) { hu = false; return su; }
        uint32_t o0 = c0, o1 = c1, o2 = c2, o3 = blk++;
        philox4x32_10(o0, o1, o2, o3, k0, k1);
        su = u53(o2, o3);
        hu = true;
        return u53(o0, o1);
    }
    // sampleGaussian; r2 == 0 is rejected too (log 0)
    RH_HD double gauss() {
#pragma clang fp contract(off)
        if (hg) { hg = false; return sg; }
        double x, y, r2;
        do {
            x = 2 * uniform() - 1;
            y = 2 * uniform() - 1;
            r2 = x * x + y * y;
        } while (r2 >= 1.0 || r2 == 0.0);
        const double fac = sqrt(-2.0 * log(r2) / r2);
        sg = y * fac;
        hg = true;
        return x * fac;
    }
};

enum InjKind { INJ_NONE = 0, INJ_SOLAR_WIND = 1, INJ_EXOSPHERE = 2, INJ_UNIFORM = 3 };

// Injector of one species for one step.  The constants of
// sampleDerivGaussian (x0 = U/vth, its mode xm, the normalisation c, d =
// (x0-xm)^2, c*xm) are computed on the host.
struct InjParams {
    int kind;
    uint32_t k0, k1, step;
    int axis, pos_wall;     // solar wind: wall normal; 1 = the +axis wall (flow towards -axis)
    double rate;            // solar wind: macroparticles per listed cell per dt
    double vth;
    double x0, xm, c, d, cxm;
    double eps;             // exosphere, uniform: the cell is shrunk by eps (epsCell)
    double R2_exobase;      // exosphere: trial positions below the exobase are discarded
    double U[3];            // uniform: bulk velocity
};

RH_HD int prob_round(double x, double r) {
    if (x <= 0) return 0;
    const int f = (int)floor(x);
    return (r < x - f) ? f + 1 : f;
}

RH_HD double deriv_gauss(RngStream& st, const InjParams& ip) {
#pragma clang fp contract(off)
    for (;;) {
        const double x = ip.xm + st.gauss();
        if (x < 0) continue;
        const double majorant = ip.cxm * exp(-0.5 * ((x - ip.xm) * (x - ip.xm) + ip.d));
        const double pdf = ip.c * x * exp(-0.5 * ((x - ip.x0) * (x - ip.x0)));
        if (st.uniform() * majorant > pdf) continue;
        return x;
    }
}

// Trial particle m of listed cell n (its own stream): position in cell units,
// velocity in m/s; false when the exosphere discards it below the exobase.
__device__ __forceinline__ bool gen_particle(const Geom& g, const InjParams& ip, int n, int i, int j, int k, uint32_t m,
                                             bool need_velocity, double (&r)[3], double (&v)[3]) {
#pragma clang fp contract(off)
    RngStream st(ip.k0, ip.k1, (uint32_t)n, ip.step, m);
    if (ip.kind == INJ_SOLAR_WIND) {
        const int a = ip.axis;
        const int l1 = a == 0 ? 1 : 0, l2 = a == 2 ? 1 : 2;
        r[a] = ip.pos_wall ? 0.0 : 1.0;
        r[l1] = st.uniform();
        r[l2] = st.uniform();
        if (!need_velocity) return true;
        const double vn = ip.vth * deriv_gauss(st, ip);
        v[a] = ip.pos_wall ? -vn : vn;
        v[l1] = ip.vth * st.gauss();
        v[l2] = ip.vth * st.gauss();
        return true;
    }
    const double dx = g.dx;
    double rm[3];
    for (int l = 0; l < 3; ++l) rm[l] = 0.5 * dx + (1.0 - ip.eps) * dx * (st.uniform() - 0.5);
    if (ip.kind == INJ_UNIFORM) {
        // InjectorUniform::injectParticles (:187-252): drifting Maxwellian
        if (!need_velocity) return true;
        for (int l = 0; l < 3; ++l) r[l] = rm[l] / dx;
        for (int l = 0; l < 3; ++l) v[l] = ip.U[l] + ip.vth * st.gauss();
        return true;
    }
    const double gx = (g.x0 + i * dx) + rm[0], gy = (g.y0 + j * dx) + rm[1], gz = (g.z0 + k * dx) + rm[2];
    if (gx * gx + gy * gy + gz * gz < ip.R2_exobase) return false;
    if (!need_velocity) return true;
    for (int l = 0; l < 3; ++l) r[l] = rm[l] / dx;
    for (int l = 0; l < 3; ++l) v[l] = ip.vth * st.gauss();
    return true;
}

// Stores the particle; returns v^2 of the stored velocity (for the inject
// kinetic energy counter, src/particle_injector.cpp:575).
template <typename Real>
__device__ __forceinline__ double store_particle(const InjParams& ip, PArrays<Real> A, uint32_t p, const double (&r)[3],
                                                 const double (&v)[3], uint8_t lc) {
    Real X = (Real)r[0], Y = (Real)r[1], Z = (Real)r[2];
    // a uniform draw just below 1 can round up to 1 in FP32 (the -axis
    // wall's coordinate is exactly 1 and stays)
    const bool sw = ip.kind == INJ_SOLAR_WIND;
    if (!sw || ip.axis != 0 || ip.pos_wall) X = X >= Real(1) ? below_one<Real>() : X;
    if (!sw || ip.axis != 1 || ip.pos_wall) Y = Y >= Real(1) ? below_one<Real>() : Y;
    if (!sw || ip.axis != 2 || ip.pos_wall) Z = Z >= Real(1) ? below_one<Real>() : Z;
    const Real vx = (Real)v[0], vy = (Real)v[1], vz = (Real)v[2];
    A.x[p] = X; A.y[p] = Y; A.z[p] = Z;
    A.vx[p] = vx; A.vy[p] = vy; A.vz[p] = vz;
    A.lc[p] = lc;
    return (double)vx * (double)vx + (double)vy * (double)vy + (double)vz * (double)vz;
}

// The new particles of listed cell n: their number (WRITE = false), or
// their state written to A[at..] (WRITE = true; indices >= cap are not
// written and set `overflow`).  Positions in cell units, velocities in m/s.
//   solar wind (initParticleCrdVel{X,Y,Z}{Pos,Neg}, :464-535): on the wall
//     face (local coordinate 0 on the +axis wall, 1 on the -axis wall),
//     uniform across it, normal speed vth * sampleDerivGaussian(U/vth)
//     towards the interior, transverse vth * sampleGaussian;
//   exosphere (InjectorExosphere::injectParticles, :1522-1591): uniform in the
//     cell shrunk by eps, discarded below the exobase, vth * sampleGaussian
//     per component.
template <typename Real, bool WRITE>
__device__ uint32_t inject_cell(const Geom& g, const InjParams& ip, int n, double rate, PArrays<Real> A,
                                uint32_t at, uint32_t cap, uint8_t lc, bool& overflow, double& v2sum) {
#pragma clang fp contract(off)
    RngStream cs(ip.k0, ip.k1, (uint32_t)n, ip.step, 0xFFFFFFFFu);
    const int N = prob_round(rate, cs.uniform());
    if (!WRITE && ip.kind == INJ_SOLAR_WIND) return (uint32_t)N;
    int i = 0, j = 0, k = 0;
    cell_ijk(g, n, i, j, k);
    uint32_t q = 0;
    for (int m = 0; m < N; ++m) {
        double r[3], v[3];
        if (!gen_particle(g, ip, n, i, j, k, (uint32_t)m, WRITE, r, v)) continue;
        if (WRITE) {
            const uint32_t p = at + q;
            if (p < cap) v2sum += store_particle(ip, A, p, r, v, lc);
            else overflow = true;
        }
        ++q;
    }
    return q;
}

// Solar wind and uniform fill: every listed cell owns `slots` = ceil(rate) fixed places past
// the store's end (cell c: slots c*slots ..), so no count and no scan are
// needed; one thread per place writes particle m of its cell if m < the
// cell's probRound draw, else marks the place removed (x = -1, skipped by
// every later pass).  The step's total goes to *total; the block's sum of
// v^2 (fixed butterfly per wave, waves in order) to v2acc[block].
template <typename Real>
__global__ __launch_bounds__(256) void k_inject_slots(Geom g, InjParams ip, int ncell, int slots,
                                                      const int* __restrict__ cells,
                                                      const uint32_t* __restrict__ np_ptr, PArrays<Real> A,
                                                      uint32_t cap, unsigned long long* __restrict__ total,
                                                      unsigned long long* __restrict__ counters,
                                                      double* __restrict__ v2acc) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    bool overflow = false, live = false;
    double v2 = 0.0;
    if (t < ncell * slots) {
        const int c = t / slots, m = t - c * slots;
        const int n = cells[c];
        RngStream cs(ip.k0, ip.k1, (uint32_t)n, ip.step, 0xFFFFFFFFu);
        const int N = prob_round(ip.rate, cs.uniform());
        const uint64_t p = (uint64_t)*np_ptr + (uint64_t)t;
        if (p >= cap) {
            overflow = m < N;
        } else if (m < N) {
            int i, j, k;
            cell_ijk(g, n, i, j, k);
            const uint8_t lc = (uint8_t)((i & (g.tx - 1)) + ((j & (g.ty - 1)) << g.ltx) +
                                         ((k & (g.tz - 1)) << (g.ltx + g.lty)));
            double r[3], v[3];
            gen_particle(g, ip, n, i, j, k, (uint32_t)m, true, r, v);
            v2 = store_particle(ip, A, (uint32_t)p, r, v, lc);
            live = true;
        } else {
            A.x[p] = Real(-1);
        }
    }
    // integer sums: the same total in any order
    const unsigned long long b = __ballot(live);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(total, (unsigned long long)__popcll(b));
    if (overflow) atomicAdd(counters + C_RANK_OVERFLOW, 1ull);  // slot 0, as every overflow site
    __shared__ double wsum[4];
    for (int off = 32; off > 0; off >>= 1) v2 += __shfl_xor(v2, off, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v2;
    __syncthreads();
    if (threadIdx.x == 0) v2acc[blockIdx.x] += ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}

// Count of new particles per listed cell (rate: per listed cell, or the
// solar wind's constant when null).
template <typename Real>
__global__ __launch_bounds__(256) void k_inject_count(Geom g, InjParams ip, int ncell, const int* __restrict__ cells,
                                                      const double* __restrict__ rate, uint32_t* __restrict__ count) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncell) return;
    bool unused = false;
    double unused_v2 = 0.0;
    PArrays<Real> none{};
    count[c] = inject_cell<Real, false>(g, ip, cells[c], rate ? rate[c] : ip.rate, none, 0, 0, 0, unused, unused_v2);
}

// The new particles of each listed cell at np + off[c] (np = the sorted
// store's size); counts the step's total into *total (integer, so the sum is
// exact), overflows of the store into the sort's overflow counter, and adds
// the cell's sum of v^2 to v2acc[c].
template <typename Real>
__global__ __launch_bounds__(256) void k_inject_fill(Geom g, InjParams ip, int ncell, const int* __restrict__ cells,
                                                     const double* __restrict__ rate, const uint32_t* __restrict__ off,
                                                     const uint32_t* __restrict__ np_ptr, PArrays<Real> A, uint32_t cap,
                                                     unsigned long long* __restrict__ total,
                                                     unsigned long long* __restrict__ counters,
                                                     double* __restrict__ v2acc) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    bool overflow = false;
    if (c < ncell) {
        const int n = cells[c];
        int i, j, k;
        cell_ijk(g, n, i, j, k);
        const uint8_t lc = (uint8_t)((i & (g.tx - 1)) + ((j & (g.ty - 1)) << g.ltx) +
                                     ((k & (g.tz - 1)) << (g.ltx + g.lty)));
        const uint64_t at = (uint64_t)*np_ptr + off[c];
        double v2 = 0.0;
        inject_cell<Real, true>(g, ip, n, rate ? rate[c] : ip.rate, A, at > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)at,
                                cap, lc, overflow, v2);
        v2acc[c] += v2;
        if (c == 0) atomicAdd(total, (unsigned long long)off[ncell]);
    }
    if (overflow) atomicAdd(counters + C_RANK_OVERFLOW, 1ull);  // slot 0
}

}  // namespace rh
