// rh_log.hpp -- device reductions behind the reference's log files
// (logWriteParticleField, src/diagnostics.h:187-816): the per-cell ion mass
// density of logCalcParticle and the field sums and extremes of logCalcField;
// and the per-cell particle moments of its VLSV output
// (src/operator_userdata.cpp:355-447).
// FP64 throughout; every sum is per thread over a fixed cell assignment, then
// a fixed butterfly per wave, a fixed order per block, and the host adds the
// block partials in block order -- the same bits every run.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

#include "rh_geom.hpp"
#include "rh_particles.hpp"

namespace rh {

// Constants of the logs (CODATA; Corsair's table is not in the reference
// tree): proton and electron mass, elementary charge, vacuum permeability and
// permittivity.
constexpr double LOG_MP = 1.672621716e-27;
constexpr double LOG_ME = 9.10938215e-31;
constexpr double LOG_E = 1.60217653e-19;
constexpr double LOG_MU0 = 4.0e-7 * 3.14159265358979323846;
constexpr double LOG_EPS0 = 8.854187817e-12;
constexpr double LOG_KB = 1.3806505e-23;

// rhoM[n] (+)= m*w * (live particles of the species in cell n); removed
// particles (x < 0) are not counted.
template <typename Real>
__global__ __launch_bounds__(256) void k_cell_mass(Geom g, const uint32_t* __restrict__ sub, PArrays<Real> A, double mw,
                                                   double* __restrict__ rhoM, int first) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= g.nc) return;
    int i, j, k;
    cell_ijk(g, n, i, j, k);
    const uint32_t kc = cell_key(g, i, j, k);
    const uint32_t p0 = sub[(size_t)kc * 8], p1 = sub[(size_t)kc * 8 + 8];
    uint32_t c = 0;
    for (uint32_t p = p0; p < p1; ++p) c += A.x[p] >= Real(0);
    rhoM[n] = (first ? 0.0 : rhoM[n]) + mw * (double)c;
}

// Per-cell bulk parameters of one species (UserDataOP::
// calcCellParticleBulkParameters, src/operator_userdata.cpp:376-447), added
// to the output population's sums.  pass 0: acc[4n..4n+3] += (sum w, sum w vx,
// sum w vy, sum w vz); pass 1 (after the population's pass 0): tacc[n] +=
// sum w |v - U|^2 with U = the population's mean velocity in the cell (0
// where the cell is empty).  Removed particles (x < 0) are skipped.
template <typename Real>
__global__ __launch_bounds__(256) void k_cell_bulk(Geom g, const uint32_t* __restrict__ sub, PArrays<Real> A, double w,
                                                   int pass, double* __restrict__ acc, double* __restrict__ tacc,
                                                   int first) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= g.nc) return;
    int i, j, k;
    cell_ijk(g, n, i, j, k);
    const uint32_t kc = cell_key(g, i, j, k);
    const uint32_t p0 = sub[(size_t)kc * 8], p1 = sub[(size_t)kc * 8 + 8];
    if (pass == 0) {
        double s0 = 0, sx = 0, sy = 0, sz = 0;
        for (uint32_t p = p0; p < p1; ++p) {
            if (A.x[p] < Real(0)) continue;
            s0 += w;
            sx += (double)A.vx[p] * w; sy += (double)A.vy[p] * w; sz += (double)A.vz[p] * w;
        }
        double* a = acc + (size_t)n * 4;
        if (first) { a[0] = s0; a[1] = sx; a[2] = sy; a[3] = sz; }
        else { a[0] += s0; a[1] += sx; a[2] += sy; a[3] += sz; }
    } else {
        const double* a = acc + (size_t)n * 4;
        const double ux = a[0] > 0.0 ? a[1] / a[0] : 0.0, uy = a[0] > 0.0 ? a[2] / a[0] : 0.0,
                     uz = a[0] > 0.0 ? a[3] / a[0] : 0.0;
        double t = 0;
        for (uint32_t p = p0; p < p1; ++p) {
            if (A.x[p] < Real(0)) continue;
            const double dx = (double)A.vx[p] - ux, dy = (double)A.vy[p] - uy, dz = (double)A.vz[p] - uz;
            t += w * (dx * dx + dy * dy + dz * dz);
        }
        tacc[n] = first ? t : tacc[n] + t;
    }
}

// The field quantities of logCalcField (:254-404), in this order: 22 sums,
// then 10 maxima, then 3 minima.
enum LogQ {
    LQ_N = 0, LQ_BX, LQ_BY, LQ_BZ, LQ_BT, LQ_DIV, LQ_B2, LQ_JX, LQ_JY, LQ_JZ, LQ_JT, LQ_EPX, LQ_EPY, LQ_EPZ, LQ_EPT,
    LQ_EP2, LQ_EX, LQ_EY, LQ_EZ, LQ_ET, LQ_E2, LQ_CE2,
    LQ_MAX_B, LQ_MAX_DIV, LQ_MAX_DIVPB, LQ_MAX_J, LQ_MAX_EP, LQ_MAX_E, LQ_MAX_DE, LQ_MAX_DI, LQ_MAX_VA, LQ_MAX_UE,
    LQ_MIN_DE, LQ_MIN_DI, LQ_MIN_TL, LQ_N_Q
};
constexpr int LQ_NSUM = LQ_MAX_B, LQ_NMAX_END = LQ_MIN_DE;
constexpr int LOG_BLOCKS = 512;

__device__ __forceinline__ double lq_combine(int q, double a, double b) {
    return q < LQ_NSUM ? a + b : (q < LQ_NMAX_END ? fmax(a, b) : fmin(a, b));
}

template <typename Real>
struct LogFields {
    V3<Real> faceB, Ji, Ep, nodeE, Ue;
    const Real* rho;
    const uint8_t* inner;
    const double* rhoM;  // Σ m w per cell (before /dV and the floor)
};

// Cells whose 26 neighbours exist, without the inner-boundary cells unless
// include_inner; partial[block * LQ_N_Q + q].
template <typename Real>
__global__ __launch_bounds__(256) void k_field_log(Geom g, LogFields<Real> f, int include_inner, double min_rhoM,
                                                   double* __restrict__ partial) {
    __shared__ double red[4][LQ_N_Q];
    double v[LQ_N_Q];
    for (int q = 0; q < LQ_N_Q; ++q) v[q] = q < LQ_NSUM ? 0.0 : (q < LQ_NMAX_END ? 0.0 : DBL_MAX);
    v[LQ_MAX_DE] = -DBL_MAX;
    v[LQ_MAX_DI] = -DBL_MAX;
    const double dV = g.dx * g.dx * g.dx;
    for (int n = blockIdx.x * 256 + threadIdx.x; n < g.nc; n += 256 * gridDim.x) {
        int i, j, k;
        cell_ijk(g, n, i, j, k);
        if (!all_nbrs(g, i, j, k)) continue;
        if (!include_inner && f.inner[n]) continue;
        const int nx = nbr(g, i, j, k, -1, 0, 0), ny = nbr(g, i, j, k, 0, -1, 0), nz = nbr(g, i, j, k, 0, 0, -1);
        const double fx = f.faceB.x[n], fy = f.faceB.y[n], fz = f.faceB.z[n];
        const double fxm = f.faceB.x[nx], fym = f.faceB.y[ny], fzm = f.faceB.z[nz];
        const double divB = fabs(((fx - fxm) + (fy - fym) + (fz - fzm)) / g.dx);
        const double Bx = 0.5 * (fx + fxm), By = 0.5 * (fy + fym), Bz = 0.5 * (fz + fzm);
        const double B2 = Bx * Bx + By * By + Bz * Bz, Bt = sqrt(B2);
        const double dpb = Bt > 0.0 ? divB / Bt : 0.0;
        v[LQ_N] += 1.0;
        v[LQ_BX] += Bx; v[LQ_BY] += By; v[LQ_BZ] += Bz; v[LQ_BT] += Bt; v[LQ_DIV] += divB; v[LQ_B2] += B2;
        v[LQ_MAX_B] = fmax(v[LQ_MAX_B], Bt);
        v[LQ_MAX_DIV] = fmax(v[LQ_MAX_DIV], divB);
        v[LQ_MAX_DIVPB] = fmax(v[LQ_MAX_DIVPB], dpb);
        const double jx = f.Ji.x[n], jy = f.Ji.y[n], jz = f.Ji.z[n];
        const double Jt = sqrt(jx * jx + jy * jy + jz * jz);
        v[LQ_JX] += jx; v[LQ_JY] += jy; v[LQ_JZ] += jz; v[LQ_JT] += Jt;
        v[LQ_MAX_J] = fmax(v[LQ_MAX_J], Jt);
        const double ex = f.Ep.x[n], ey = f.Ep.y[n], ez = f.Ep.z[n];
        const double Ep2 = ex * ex + ey * ey + ez * ez, Ept = sqrt(Ep2);
        v[LQ_EPX] += ex; v[LQ_EPY] += ey; v[LQ_EPZ] += ez; v[LQ_EPT] += Ept; v[LQ_EP2] += Ep2;
        v[LQ_MAX_EP] = fmax(v[LQ_MAX_EP], Ept);
        const double Ex = f.nodeE.x[n], Ey = f.nodeE.y[n], Ez = f.nodeE.z[n];
        const double E2 = Ex * Ex + Ey * Ey + Ez * Ez, Et = sqrt(E2);
        v[LQ_EX] += Ex; v[LQ_EY] += Ey; v[LQ_EZ] += Ez; v[LQ_ET] += Et; v[LQ_E2] += E2;
        v[LQ_MAX_E] = fmax(v[LQ_MAX_E], Et);
        const double ux = f.Ue.x[n], uy = f.Ue.y[n], uz = f.Ue.z[n];
        const double cEx = -(uy * Bz - uz * By), cEy = -(uz * Bx - ux * Bz), cEz = -(ux * By - uy * Bx);
        v[LQ_CE2] += cEx * cEx + cEy * cEy + cEz * cEz;
        v[LQ_MAX_UE] = fmax(v[LQ_MAX_UE], sqrt(ux * ux + uy * uy + uz * uz));
        const double rq = f.rho[n];
        if (rq > 0.0) {
            const double de = sqrt(LOG_ME / (LOG_MU0 * LOG_E * rq));
            v[LQ_MAX_DE] = fmax(v[LQ_MAX_DE], de);
            v[LQ_MIN_DE] = fmin(v[LQ_MIN_DE], de);
        }
        if (Bt > 0.0) v[LQ_MIN_TL] = fmin(v[LQ_MIN_TL], 2.0 * 3.14159265358979323846 * LOG_MP / (LOG_E * Bt));
        double rm = f.rhoM[n] / dV;
        if (rm < min_rhoM) rm = min_rhoM;
        if (rm > 0.0) {
            const double di = sqrt(LOG_MP * LOG_MP / (LOG_MU0 * LOG_E * LOG_E * rm));
            v[LQ_MAX_DI] = fmax(v[LQ_MAX_DI], di);
            v[LQ_MIN_DI] = fmin(v[LQ_MIN_DI], di);
            v[LQ_MAX_VA] = fmax(v[LQ_MAX_VA], Bt / sqrt(LOG_MU0 * rm));
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int q = 0; q < LQ_N_Q; ++q) {
        double a = v[q];
        for (int off = 32; off > 0; off >>= 1) a = lq_combine(q, a, __shfl_xor(a, off, 64));
        if (lane == 0) red[w][q] = a;
    }
    __syncthreads();
    if (threadIdx.x < LQ_N_Q) {
        const int q = threadIdx.x;
        double a = red[0][q];
        for (int ww = 1; ww < 4; ++ww) a = lq_combine(q, a, red[ww][q]);
        partial[(size_t)blockIdx.x * LQ_N_Q + q] = a;
    }
}

// Column sums of a rows x cols FP64 array (cols <= 4): partial[block*cols + c].
__global__ __launch_bounds__(256) void k_sum_cols(const double* __restrict__ a, size_t rows, int cols,
                                                 double* __restrict__ partial) {
    __shared__ double red[4][4];
    double s[4] = {0, 0, 0, 0};
    for (size_t r = blockIdx.x * 256 + threadIdx.x; r < rows; r += 256 * (size_t)gridDim.x)
        for (int c = 0; c < cols; ++c) s[c] += a[r * cols + c];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int c = 0; c < cols; ++c) {
        double x = s[c];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) red[w][c] = x;
    }
    __syncthreads();
    if ((int)threadIdx.x < cols)
        partial[(size_t)blockIdx.x * cols + threadIdx.x] =
            ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

}  // namespace rh
