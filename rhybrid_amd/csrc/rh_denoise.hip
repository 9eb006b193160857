// rh_denoise.hip -- the learned moment denoiser on gfx950 matrix cores.
//
// Four 3x3x3 convolutions (7->32->32->32->4 channels, no biases, ReLU after
// the first three, residual on n and J).  Every convolution -- forward,
// backward-data and backward-weights -- is an implicit GEMM on
// v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32 accumulation):
//
//  * forward / backward-data (k_conv): M = 16 output channels, N = 16 output
//    voxels along x, K = 32 = one tap's 32 input channels (or four taps of 8
//    channels for the 8-channel layers).  A workgroup computes a 16x4x4
//    voxel tile (one wave per z plane, four N tiles of one x row each); the
//    18x6x6 input halo sits in LDS in channels-last order, so a lane's voxel
//    fragment (8 channels of one voxel) is one 16-byte LDS read and a wave
//    reads 1 KB contiguously; each input row's fragment serves the (up to)
//    three output rows it reaches through the dy taps.  Weight fragments are
//    pre-packed in the lane order of the instruction and read from L1/L2.
//    With the weights as the A operand a lane ends with four consecutive
//    channels of one voxel, stored as one vector.  The epilogue applies ReLU
//    (forward), the ReLU mask of the forward activation (backward), or the
//    FP32 residual (last layer).
//  * backward-weights (k_wgrad): dW[tap][ci][co] = sum over voxels of
//    in[v + tap][ci] * G[v][co]: M = 16 input channels, N = 16 output
//    channels, K = 32 voxels.  The workgroups are persistent; each stages a
//    16x4x2 voxel tile transposed into LDS (channel-major, three copies of
//    the input shifted by dx = -1, 0, +1 so that every 8-voxel K run is one
//    aligned 16-byte read) and keeps its taps' partial sums in registers;
//    k_wgrad_reduce adds the per-workgroup partials in a fixed order.
//
// Fragment lane layouts (checked on the device): for 16x16x32, lane l holds
// A[l%16][8*(l/16) .. +7], B[8*(l/16) .. +7][l%16] and D[4*(l/16) + i][l%16].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rh_denoise.h"

namespace rhdn {

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int TX = 16, TY = 4, TZ = 4;             // forward output tile
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2;
constexpr int HV = HX * HY * HZ;
constexpr int NTAP = 27;

struct Layer {
    int cin, cout, offset;
};
constexpr Layer LAYERS[4] = {{7, 32, 0}, {32, 32, 6048}, {32, 32, 33696}, {32, 4, 61344}};

enum { EP_RELU = 0, EP_MASK = 1, EP_RESID = 2 };

struct ConvArgs {
    const bf16* in;          // [V][CIN] channels-last
    const bf16x8* wf;        // packed B fragments
    int X, Y, Z, nzb;        // volume and z tiles per volume
    bf16* out;               // EP_RELU / EP_MASK: [V][NT*16]
    const bf16* mask;        // EP_MASK: forward activation [V][NT*16]
    float* yout;             // EP_RESID: [V][4]
    const float* xres;       // EP_RESID: physical input [V][7]
    float s[4];              // EP_RESID: output scales
};

__device__ __forceinline__ floatx4 mfma(bf16x8 a, bf16x8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Convolution over zero-padded volumes; CIN in {8, 32}, NT output tiles of 16.
// A workgroup computes a 16x4xTZ_ tile, one wave per z plane (TZ_ = 4 or 8:
// the taller tile stages 10 halo planes for 8 output planes instead of 6 for 4).
template <int CIN, int NT, int EP, int TZ_>
__global__ __launch_bounds__(64 * TZ_) void k_conv(ConvArgs a) {
    constexpr int CH = CIN / 8;  // 16-byte chunks per voxel
    constexpr int HZ_ = TZ_ + 2, HV_ = HX * HY * HZ_, NTH = 64 * TZ_;
    __shared__ bf16x8 tile[HV_ * CH];
    // the first weight group of the 32-channel path, loaded while the halo is staged
    bf16x8 bnx[3][NT];
    if constexpr (CIN == 32) {
        const int l = threadIdx.x & 63;
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) bnx[dy][nt] = a.wf[((dy * 3) * NT + nt) * 64 + l];
    }
    const int bx = blockIdx.x * TX, by = blockIdx.y * TY;
    const int p = blockIdx.z / a.nzb, bz = (blockIdx.z % a.nzb) * TZ_;
    const size_t vol = (size_t)a.X * a.Y * a.Z;
    const bf16* inp = a.in + (size_t)p * vol * CIN;
    // halo staging: all of a thread's loads are issued before its LDS stores
    constexpr int NST = (HV_ * CH + NTH - 1) / NTH;
    bf16x8 stage[NST];
#pragma unroll
    for (int q = 0; q < NST; ++q) {
        const int idx = (int)threadIdx.x + NTH * q;
        const int v = idx / CH, c8 = idx - v * CH;
        const int hx = v % HX, hy = (v / HX) % HY, hz = v / (HX * HY);
        const int gx = bx + hx - 1, gy = by + hy - 1, gz = bz + hz - 1;
#pragma unroll
        for (int j = 0; j < 8; ++j) stage[q][j] = (bf16)0.0f;
        if (idx < HV_ * CH && gx >= 0 && gx < a.X && gy >= 0 && gy < a.Y && gz >= 0 && gz < a.Z)
            stage[q] = *reinterpret_cast<const bf16x8*>(inp + (((size_t)gz * a.Y + gy) * a.X + gx) * CIN + c8 * 8);
    }
#pragma unroll
    for (int q = 0; q < NST; ++q) {
        const int idx = (int)threadIdx.x + NTH * q;
        if (idx < HV_ * CH) tile[idx] = stage[q];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m = lane & 15, kg = lane >> 4;
    floatx4 acc[TY][NT];
#pragma unroll
    for (int my = 0; my < TY; ++my)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[my][nt] = floatx4{0.f, 0.f, 0.f, 0.f};
    if constexpr (CIN == 32) {
        // Input row r of the halo (for taps dz, dx) feeds output rows my with
        // dy = r - my in 0..2: each A fragment is read from LDS once and used
        // with the weights of its (up to) three dy taps.  The weights of the
        // next (dz, dx) group are loaded while the current group computes.
#pragma unroll
        for (int grp = 0; grp < 9; ++grp) {
            const int dz = grp / 3, dx = grp % 3;
            bf16x8 b[3][NT];
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) b[dy][nt] = bnx[dy][nt];
            if (grp + 1 < 9) {
                const int dz1 = (grp + 1) / 3, dx1 = (grp + 1) % 3;
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                        bnx[dy][nt] = a.wf[((dz1 * 9 + dy * 3 + dx1) * NT + nt) * 64 + lane];
            }
#pragma unroll
            for (int r = 0; r < TY + 2; ++r) {
                const int v = ((w + dz) * HY + r) * HX + (m + dx);
                const bf16x8 av = tile[v * CH + kg];
#pragma unroll
                for (int my = 0; my < TY; ++my) {
                    const int dy = r - my;
                    if (dy < 0 || dy > 2) continue;
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) acc[my][nt] = mfma(b[dy][nt], av, acc[my][nt]);
                }
            }
        }
    } else {
        // four taps per instruction: lane group kg takes tap 4g+kg (the
        // weights of the padding tap 27 are zero)
#pragma unroll
        for (int g = 0; g < 7; ++g) {
            int tap = 4 * g + kg;
            tap = tap < NTAP ? tap : NTAP - 1;
            const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
            bf16x8 b[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) b[nt] = a.wf[(g * NT + nt) * 64 + lane];
#pragma unroll
            for (int my = 0; my < TY; ++my) {
                const int v = ((w + dz) * HY + (my + dy)) * HX + (m + dx);
                const bf16x8 av = tile[v];
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) acc[my][nt] = mfma(b[nt], av, acc[my][nt]);
            }
        }
    }
    // The weights are the A operand and the voxels the B operand, so lane
    // (m, kg) holds D[channel nt*16 + 4*kg + i][voxel x = m] of row (my, w):
    // four consecutive channels of one voxel, stored as one 8-byte vector
    // (the FP32 residual of the last layer: channels 0-3 as one 16-byte one).
    const int z = bz + w;
    const int x = bx + m;
#pragma unroll
    for (int my = 0; my < TY; ++my) {
        const int y = by + my;
        const size_t v = (size_t)p * vol + ((size_t)z * a.Y + y) * a.X + x;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int c0 = nt * 16 + 4 * kg;
            if constexpr (EP == EP_RELU) {
                bf16x4 o;
#pragma unroll
                for (int i = 0; i < 4; ++i) o[i] = (bf16)fmaxf(acc[my][nt][i], 0.0f);
                *reinterpret_cast<bf16x4*>(a.out + v * (NT * 16) + c0) = o;
            } else if constexpr (EP == EP_MASK) {
                const bf16x4 mk = *reinterpret_cast<const bf16x4*>(a.mask + v * (NT * 16) + c0);
                bf16x4 o;
#pragma unroll
                for (int i = 0; i < 4; ++i) o[i] = (bf16)(acc[my][nt][i] * ((float)mk[i] > 0.0f ? 1.0f : 0.0f));
                *reinterpret_cast<bf16x4*>(a.out + v * (NT * 16) + c0) = o;
            } else {
                if (nt == 0 && kg == 0) {
                    floatx4 o;
#pragma unroll
                    for (int i = 0; i < 4; ++i) o[i] = a.xres[v * 7 + i] + a.s[i] * acc[my][nt][i];
                    *reinterpret_cast<floatx4*>(a.yout + v * 4) = o;
                }
            }
        }
    }
}

// Backward-weights.  CIN, COUT = channels of the layer input and of G
// (8 or 32); tiles of 16x4x2 output voxels; each wave owns the taps
// w, w+4, ... and all (ci, co) tiles of them.
constexpr int WX = 16, WY = 4, WZ = 2, WK = WX * WY * WZ;  // 128 voxels = 4 K steps
constexpr int WHY = WY + 2, WHZ = WZ + 2;
constexpr int WTAPS = 7;  // taps per wave (27 = 7+7+7+6)

template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_wgrad(const bf16* __restrict__ in, const bf16* __restrict__ G, int X,
                                               int Y, int Z, int nvol, float* __restrict__ part) {
    constexpr int CI = CIN < 16 ? 16 : CIN;   // M extent (padded to a tile)
    constexpr int CO = COUT < 16 ? 16 : COUT; // N extent
    constexpr int MT = CI / 16, NT = CO / 16;
    // 16-byte fragment reads: both arrays aligned to 16 bytes
    __shared__ __attribute__((aligned(16))) bf16 Xs[3 * CI * WHZ * WHY * WX];  // [dx][ci][z][y][x]
    __shared__ __attribute__((aligned(16))) bf16 Gt[CO * WK];                  // [co][voxel]
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m = lane & 15, kg = lane >> 4;
    const int ntx = X / WX, nty = Y / WY, ntz = Z / WZ;
    const int ntiles = ntx * nty * ntz * nvol;
    const size_t vol = (size_t)X * Y * Z;
    floatx4 acc[WTAPS][MT][NT];
#pragma unroll
    for (int t = 0; t < WTAPS; ++t)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[t][mt][nt] = floatx4{0.f, 0.f, 0.f, 0.f};
    // zero the padding channels once (never written by the staging below)
    for (int idx = threadIdx.x; idx < 3 * CI * WHZ * WHY * WX; idx += blockDim.x) Xs[idx] = (bf16)0.0f;
    for (int idx = threadIdx.x; idx < CO * WK; idx += blockDim.x) Gt[idx] = (bf16)0.0f;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        int r = tile;
        const int tx = r % ntx; r /= ntx;
        const int ty = r % nty; r /= nty;
        const int tz = r % ntz; r /= ntz;
        const int p = r;
        const int bx = tx * WX, by = ty * WY, bz = tz * WZ;
        __syncthreads();  // the previous tile's reads are done
        // input halo: 18 x 6 x 4 voxels, each written to the three shifted copies
        constexpr int CHI = CIN / 8;
        for (int idx = threadIdx.x; idx < (WX + 2) * WHY * WHZ * CHI; idx += blockDim.x) {
            const int c8 = idx % CHI;
            const int v = idx / CHI;
            const int hx = v % (WX + 2), hy = (v / (WX + 2)) % WHY, hz = v / ((WX + 2) * WHY);
            const int gx = bx + hx - 1, gy = by + hy - 1, gz = bz + hz - 1;
            bf16x8 val;
#pragma unroll
            for (int j = 0; j < 8; ++j) val[j] = (bf16)0.0f;
            if (gx >= 0 && gx < X && gy >= 0 && gy < Y && gz >= 0 && gz < Z)
                val = *reinterpret_cast<const bf16x8*>(in + ((size_t)p * vol + ((size_t)gz * Y + gy) * X + gx) * CIN +
                                                       c8 * 8);
            // copy dx holds in[x + dx - 1] at column x: column = hx - dx
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int col = hx - dx;
                if (col < 0 || col >= WX) continue;
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    Xs[(((dx * CI + c8 * 8 + j) * WHZ + hz) * WHY + hy) * WX + col] = val[j];
            }
        }
        constexpr int CHG = COUT / 8;
        for (int idx = threadIdx.x; idx < WK * CHG; idx += blockDim.x) {
            const int c8 = idx % CHG, k = idx / CHG;
            const int x = k % WX, y = (k / WX) % WY, zz = k / (WX * WY);
            const bf16x8 val = *reinterpret_cast<const bf16x8*>(
                G + ((size_t)p * vol + ((size_t)(bz + zz) * Y + by + y) * X + bx + x) * COUT + c8 * 8);
#pragma unroll
            for (int j = 0; j < 8; ++j) Gt[(c8 * 8 + j) * WK + k] = val[j];
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < WK / 32; ++ks) {
            // this lane's 8 voxels: row r = 2*ks + kg/2 of the tile, columns xo..xo+7
            const int row = 2 * ks + (kg >> 1), xo = (kg & 1) * 8;
            const int yy = row % WY, zz = row / WY;
            bf16x8 b[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                b[nt] = *reinterpret_cast<const bf16x8*>(&Gt[(nt * 16 + m) * WK + ks * 32 + kg * 8]);
#pragma unroll
            for (int t = 0; t < WTAPS; ++t) {
                const int tap = w + 4 * t;
                if (tap >= NTAP) break;
                const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    const bf16x8 av = *reinterpret_cast<const bf16x8*>(
                        &Xs[(((dx * CI + mt * 16 + m) * WHZ + zz + dz) * WHY + yy + dy) * WX + xo]);
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt) acc[t][mt][nt] = mfma(av, b[nt], acc[t][mt][nt]);
                }
            }
        }
    }
    // partial[block][tap][ci][co]; lane holds D[ci = 4*kg + i][co = m] of each tile
    float* out = part + (size_t)blockIdx.x * NTAP * CI * CO;
#pragma unroll
    for (int t = 0; t < WTAPS; ++t) {
        const int tap = w + 4 * t;
        if (tap >= NTAP) break;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    out[((size_t)tap * CI + mt * 16 + 4 * kg + i) * CO + nt * 16 + m] = acc[t][mt][nt][i];
    }
}

// grad[offset + (co*cin + ci)*27 + tap] = sum over workgroups in a fixed
// order: a block of 256 threads takes 8 consecutive weights (co fastest, the
// partials' memory order) x 32 groups of workgroups; group g sums workgroups
// g, g+32, ... and the 32 group sums are added in order of g.
constexpr int RED_W = 8, RED_G = 32;
__global__ __launch_bounds__(RED_W * RED_G) void k_wgrad_reduce(const float* __restrict__ part, int nblk, int CI,
                                                                int CO, int cin, int cout, float* __restrict__ grad) {
    __shared__ double red[RED_G][RED_W];
    const int wl = threadIdx.x % RED_W, gi = threadIdx.x / RED_W;
    const int e = blockIdx.x * RED_W + wl;
    const int ne = cout * cin * NTAP;
    const int co = e % cout, ci = (e / cout) % cin, tap = e / (cout * cin);
    const size_t stride = (size_t)NTAP * CI * CO;
    const size_t off = ((size_t)tap * CI + ci) * CO + co;
    double s = 0.0;
    if (e < ne)
        for (int b = gi; b < nblk; b += RED_G) s += part[b * stride + off];
    red[gi][wl] = s;
    __syncthreads();
    if (gi == 0 && e < ne) {
        double t = 0.0;
        for (int g = 0; g < RED_G; ++g) t += red[g][wl];
        grad[((size_t)co * cin + ci) * NTAP + tap] = (float)t;
    }
}

// packed B fragments of one layer.  KP = 32: fragment f = tap*NT + nt,
// lane l, element j -> (co = nt*16 + l%16, ci = 8*(l/16) + j, tap).  KP = 8:
// f = g*NT + nt -> (co = nt*16 + l%16, ci = j, tap = 4g + l/16).  flip: the
// transposed, spatially flipped weights of the backward-data convolution.
__global__ void k_pack(const float* __restrict__ W, int cin, int cout, int flip, int KP, int NT, int nfrag,
                       bf16x8* __restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nfrag * 64) return;
    const int lane = e & 63, f = e >> 6;
    const int nt = f % NT, blk = f / NT;
    const int co = nt * 16 + (lane & 15);
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        int ci, tap;
        if (KP == 32) { ci = 8 * (lane >> 4) + j; tap = blk; }
        else { ci = j; tap = 4 * blk + (lane >> 4); }
        // the convolution's own (co, ci) ranges: the flipped one swaps them
        const int oc = flip ? cin : cout, ic = flip ? cout : cin;
        float x = 0.0f;
        if (co < oc && ci < ic && tap < NTAP) {
            x = flip ? W[((size_t)ci * cin + co) * NTAP + (NTAP - 1 - tap)] : W[((size_t)co * cin + ci) * NTAP + tap];
        }
        v[j] = (bf16)x;
    }
    out[e] = v;
}

// FP32 [V][7] -> BF16 [V][8] normalised (channel 7 = 0)
__global__ void k_pack_input(const float* __restrict__ x, size_t V, float i0, float i1, float i2, float i3,
                             float i4, float i5, float i6, bf16x8* __restrict__ out) {
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const float* s = x + v * 7;
    bf16x8 o;
    o[0] = (bf16)(s[0] * i0); o[1] = (bf16)(s[1] * i1); o[2] = (bf16)(s[2] * i2); o[3] = (bf16)(s[3] * i3);
    o[4] = (bf16)(s[4] * i4); o[5] = (bf16)(s[5] * i5); o[6] = (bf16)(s[6] * i6); o[7] = (bf16)0.0f;
    out[v] = o;
}

// normalised difference d = (y - t)/s; loss partial sums of d^2 per block
// (FP64, fixed order); G = 2 d / (4 V) as BF16 [V][8]
constexpr int LOSS_THREADS = 256;
__global__ __launch_bounds__(LOSS_THREADS) void k_loss(const float* __restrict__ y, const float* __restrict__ t,
                                                       size_t V, float i0, float i1, float i2, float i3,
                                                       float gscale, bf16x8* __restrict__ G,
                                                       double* __restrict__ part) {
    __shared__ double red[LOSS_THREADS];
    double acc = 0.0;
    const float is[4] = {i0, i1, i2, i3};
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (size_t)gridDim.x * blockDim.x) {
        bf16x8 g;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float d = (y[v * 4 + c] - t[v * 4 + c]) * is[c];
            acc += (double)d * d;
            g[c] = (bf16)(gscale * d);
        }
#pragma unroll
        for (int c = 4; c < 8; ++c) g[c] = (bf16)0.0f;
        G[v] = g;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = LOSS_THREADS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// one block of LOSS_THREADS: strided partial sums, then a fixed-order tree
__global__ __launch_bounds__(LOSS_THREADS) void k_sum(const double* __restrict__ part, int n, double scale,
                                                      float* __restrict__ out) {
    __shared__ double red[LOSS_THREADS];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += LOSS_THREADS) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = LOSS_THREADS / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = (float)(red[0] * scale);
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

}  // namespace rhdn

// --------------------------------------------------------------- host side

using namespace rhdn;

namespace {

struct HipErr {
    std::string msg;
};
struct ArgErr {
    std::string msg;
};
struct CapErr {
    std::string msg;
};

#define DN_HIP(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) throw HipErr{std::string(#x) + ": " + hipGetErrorString(e_)};         \
    } while (0)

template <typename T>
struct Buf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count) {
        if (p) DN_HIP(hipFree(p));
        p = nullptr;
        n = count;
        if (count) DN_HIP(hipMalloc((void**)&p, count * sizeof(T)));
    }
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
};

unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

// Scratch of rh_dn_forward_static: the packed input volume (8 BF16 channels
// per voxel) and two 32-channel BF16 activation volumes, for at most
// RH_DN_STATIC_MAX_VOXELS voxels (144 B per voxel, 36 MiB).  Every GPU the
// library runs on holds its own copy until the process ends.  In host code
// the name is only a stand-in for the symbol: the device address is what
// hipGetSymbolAddress returns with that device current (scratch_on).
constexpr size_t STATIC_X8_BYTES = (size_t)RH_DN_STATIC_MAX_VOXELS * 8 * sizeof(bf16);
constexpr size_t STATIC_ACT_BYTES = (size_t)RH_DN_STATIC_MAX_VOXELS * 32 * sizeof(bf16);
static __device__ __align__(256) unsigned char scratch[STATIC_X8_BYTES + 2 * STATIC_ACT_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_dev[MAX_DEVICES] = {};

// The address of the scratch on device `dev`, which must be the current
// device: looked up once per device and kept.
void* scratch_on(int dev) {
    if (dev < 0 || dev >= MAX_DEVICES) throw ArgErr{"device index out of range for the static scratch"};
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        DN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw HipErr{"hipGetSymbolAddress gave no address for the static scratch"};
        scratch_dev[dev] = p;
    }
    return scratch_dev[dev];
}

}  // namespace

struct rh_denoiser {
    hipStream_t stream = nullptr;      // where the work is ordered
    hipStream_t own_stream = nullptr;  // the denoiser's own stream
    std::string err;
    int64_t max_voxels = 0;
    float scale[7] = {1, 1, 1, 1, 1, 1, 1};
    Buf<float> w, g, m1, m2, loss;
    Buf<bf16x8> fw[4], bw[4];  // forward fragments, backward-data fragments (layers 2..4)
    Buf<bf16> x8, a1, a2, a3, g4, g3, g2, g1;
    Buf<float> y;
    Buf<float> part;
    Buf<double> lpart;
    int adam_step = 0;
    int wgrad_blocks = 512;  // persistent workgroups (two rounds of one per CU)
    int device = 0;                  // the device current at creation (own_stream's)
    unsigned char* static_scratch = nullptr;  // that device's copy of `scratch`
    // timing
    bool timing = false;
    bool no_events = false;  // rh_dn_forward_static: launches record no events
    std::map<std::string, std::pair<double, int64_t>> times;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;

    template <typename K, typename... Args>
    void launch(const char* name, K kernel, dim3 grid, dim3 block, Args... args) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        const bool timed = timing && !no_events;
        if (timed) {
            DN_HIP(hipEventCreate(&e0));
            DN_HIP(hipEventCreate(&e1));
            DN_HIP(hipEventRecord(e0, stream));
        }
        hipLaunchKernelGGL(kernel, grid, block, 0, stream, args...);
        DN_HIP(hipGetLastError());
        if (timed) {
            DN_HIP(hipEventRecord(e1, stream));
            pending.push_back({name, {e0, e1}});
        }
    }
    void collect() {
        for (auto& pe : pending) {
            DN_HIP(hipEventSynchronize(pe.second.second));
            float ms = 0;
            DN_HIP(hipEventElapsedTime(&ms, pe.second.first, pe.second.second));
            auto& t = times[pe.first];
            t.first += ms;
            t.second += 1;
            (void)hipEventDestroy(pe.second.first);
            (void)hipEventDestroy(pe.second.second);
        }
        pending.clear();
    }

    static int nfrag(int kp, int nt) { return (kp == 32 ? NTAP : 7) * nt; }

    void alloc(int64_t maxv) {
        max_voxels = maxv;
        w.alloc(RH_DN_N_WEIGHTS);
        g.alloc(RH_DN_N_WEIGHTS);
        m1.alloc(RH_DN_N_WEIGHTS);
        m2.alloc(RH_DN_N_WEIGHTS);
        loss.alloc(1);
        // forward: L1 (KP 8, NT 2), L2, L3 (KP 32, NT 2), L4 (KP 32, NT 1)
        fw[0].alloc((size_t)nfrag(8, 2) * 64);
        fw[1].alloc((size_t)nfrag(32, 2) * 64);
        fw[2].alloc((size_t)nfrag(32, 2) * 64);
        fw[3].alloc((size_t)nfrag(32, 1) * 64);
        // backward-data of L2, L3 (32 -> 32) and L4 (4 -> 32, KP 8)
        bw[1].alloc((size_t)nfrag(32, 2) * 64);
        bw[2].alloc((size_t)nfrag(32, 2) * 64);
        bw[3].alloc((size_t)nfrag(8, 2) * 64);
        const size_t V = (size_t)maxv;
        x8.alloc(V * 8);
        a1.alloc(V * 32);
        a2.alloc(V * 32);
        a3.alloc(V * 32);
        g4.alloc(V * 8);
        g3.alloc(V * 32);
        g2.alloc(V * 32);
        g1.alloc(V * 32);
        y.alloc(V * 4);
        part.alloc((size_t)wgrad_blocks * NTAP * 32 * 32);
        lpart.alloc(1024);
        DN_HIP(hipMemsetAsync(w.p, 0, RH_DN_N_WEIGHTS * sizeof(float), stream));
        reset_adam();
        pack();
    }
    void reset_adam() {
        DN_HIP(hipMemsetAsync(m1.p, 0, RH_DN_N_WEIGHTS * sizeof(float), stream));
        DN_HIP(hipMemsetAsync(m2.p, 0, RH_DN_N_WEIGHTS * sizeof(float), stream));
        adam_step = 0;
    }
    void pack() {
        const float* W = w.p;
        auto pk = [&](int l, int flip, int kp, int nt, Buf<bf16x8>& out) {
            const int nf = nfrag(kp, nt);
            launch("pack", k_pack, dim3(cdiv((size_t)nf * 64, 256)), dim3(256), W + LAYERS[l].offset, LAYERS[l].cin,
                   LAYERS[l].cout, flip, kp, nt, nf, out.p);
        };
        pk(0, 0, 8, 2, fw[0]);
        pk(1, 0, 32, 2, fw[1]);
        pk(2, 0, 32, 2, fw[2]);
        pk(3, 0, 32, 1, fw[3]);
        pk(1, 1, 32, 2, bw[1]);
        pk(2, 1, 32, 2, bw[2]);
        pk(3, 1, 8, 2, bw[3]);
    }

    static void check_shape(int P, int X, int Y, int Z) {
        if (P <= 0 || X <= 0 || Y <= 0 || Z <= 0) throw ArgErr{"volume dimensions must be positive"};
        if (X % 16 || Y % 4 || Z % 4) throw ArgErr{"X must be a multiple of 16, Y and Z multiples of 4"};
    }
    void check_dims(int P, int X, int Y, int Z) const {
        check_shape(P, X, Y, Z);
        if ((int64_t)P * X * Y * Z > max_voxels) throw CapErr{"P*X*Y*Z exceeds the voxel capacity given at creation"};
    }

    template <int CIN, int NT, int EP>
    void conv(const char* name, const bf16* in, const bf16x8* wf, int P, int X, int Y, int Z, bf16* out,
              const bf16* mask, float* yo, const float* xres) {
        ConvArgs a{};
        a.in = in;
        a.wf = wf;
        a.X = X;
        a.Y = Y;
        a.Z = Z;
        a.out = out;
        a.mask = mask;
        a.yout = yo;
        a.xres = xres;
        for (int c = 0; c < 4; ++c) a.s[c] = scale[c];
        if (Z % 8 == 0) {
            a.nzb = Z / 8;
            launch(name, k_conv<CIN, NT, EP, 8>, dim3(X / TX, Y / TY, P * a.nzb), dim3(512), a);
        } else {
            a.nzb = Z / 4;
            launch(name, k_conv<CIN, NT, EP, 4>, dim3(X / TX, Y / TY, P * a.nzb), dim3(256), a);
        }
    }

    // The forward pass with the packed input in bx8 and the activations of
    // layers 1-3 in b1, b2, b3 (b3 may be b1: layer 3 reads only b2).
    void forward_with(const float* x, float* yo, int P, int X, int Y, int Z, bf16* bx8, bf16* b1, bf16* b2,
                      bf16* b3) {
        const size_t V = (size_t)P * X * Y * Z;
        launch("pack_input", k_pack_input, dim3(cdiv(V, 256)), dim3(256), x, V, 1.0f / scale[0], 1.0f / scale[1],
               1.0f / scale[2], 1.0f / scale[3], 1.0f / scale[4], 1.0f / scale[5], 1.0f / scale[6],
               reinterpret_cast<bf16x8*>(bx8));
        conv<8, 2, EP_RELU>("conv1", bx8, fw[0].p, P, X, Y, Z, b1, nullptr, nullptr, nullptr);
        conv<32, 2, EP_RELU>("conv2", b1, fw[1].p, P, X, Y, Z, b2, nullptr, nullptr, nullptr);
        conv<32, 2, EP_RELU>("conv3", b2, fw[2].p, P, X, Y, Z, b3, nullptr, nullptr, nullptr);
        conv<32, 1, EP_RESID>("conv4", b3, fw[3].p, P, X, Y, Z, nullptr, nullptr, yo, x);
    }

    // the backward pass reads all three activations: separate buffers
    void forward(const float* x, float* yo, int P, int X, int Y, int Z) {
        forward_with(x, yo, P, X, Y, Z, x8.p, a1.p, a2.p, a3.p);
    }

    template <int CIN, int COUT>
    void wgrad(const char* name, const bf16* in, const bf16* G, int P, int X, int Y, int Z, int layer) {
        constexpr int CI = CIN < 16 ? 16 : CIN, CO = COUT < 16 ? 16 : COUT;
        const int ntiles = (X / WX) * (Y / WY) * (Z / WZ) * P;
        const int nblk = ntiles < wgrad_blocks ? ntiles : wgrad_blocks;
        launch(name, k_wgrad<CIN, COUT>, dim3(nblk), dim3(256), in, G, X, Y, Z, P, part.p);
        const Layer& L = LAYERS[layer];
        launch("wgrad_reduce", k_wgrad_reduce, dim3(cdiv((size_t)L.cin * L.cout * NTAP, RED_W)), dim3(RED_W * RED_G),
               (const float*)part.p, nblk, CI, CO, L.cin, L.cout, g.p + L.offset);
    }

    // forward + loss + backward into g; loss left in `loss`
    void gradient(const float* x, const float* t, int P, int X, int Y, int Z) {
        const size_t V = (size_t)P * X * Y * Z;
        forward(x, y.p, P, X, Y, Z);
        const int lb = (int)std::min<size_t>(1024, cdiv(V, LOSS_THREADS));
        const float gscale = (float)(2.0 / (4.0 * (double)V));
        launch("loss", k_loss, dim3(lb), dim3(LOSS_THREADS), (const float*)y.p, t, V, 1.0f / scale[0],
               1.0f / scale[1], 1.0f / scale[2], 1.0f / scale[3], gscale, reinterpret_cast<bf16x8*>(g4.p), lpart.p);
        launch("loss_sum", k_sum, dim3(1), dim3(LOSS_THREADS), (const double*)lpart.p, lb, 1.0 / (4.0 * (double)V), loss.p);
        // backward-data with the ReLU masks of the forward activations
        conv<8, 2, EP_MASK>("dgrad4", g4.p, bw[3].p, P, X, Y, Z, g3.p, a3.p, nullptr, nullptr);
        conv<32, 2, EP_MASK>("dgrad3", g3.p, bw[2].p, P, X, Y, Z, g2.p, a2.p, nullptr, nullptr);
        conv<32, 2, EP_MASK>("dgrad2", g2.p, bw[1].p, P, X, Y, Z, g1.p, a1.p, nullptr, nullptr);
        // backward-weights
        wgrad<32, 8>("wgrad4", a3.p, g4.p, P, X, Y, Z, 3);
        wgrad<32, 32>("wgrad3", a2.p, g3.p, P, X, Y, Z, 2);
        wgrad<32, 32>("wgrad2", a1.p, g2.p, P, X, Y, Z, 1);
        wgrad<8, 32>("wgrad1", x8.p, g1.p, P, X, Y, Z, 0);
    }

    void adam(float lr) {
        ++adam_step;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, (float)adam_step), c2 = 1.0f - std::pow(b2, (float)adam_step);
        launch("adam", k_adam, dim3(cdiv(RH_DN_N_WEIGHTS, 256)), dim3(256), w.p, (const float*)g.p, m1.p, m2.p,
               (int)RH_DN_N_WEIGHTS, lr, b1, b2, 1e-8f, c1, c2);
        pack();
    }
};

namespace {

template <typename F>
int guarded(rh_denoiser* d, F&& f) {
    try {
        f();
        return RH_DN_OK;
    } catch (const ArgErr& e) {
        if (d) d->err = e.msg;
        return RH_DN_ERR_ARG;
    } catch (const CapErr& e) {
        if (d) d->err = e.msg;
        return RH_DN_ERR_CAPACITY;
    } catch (const HipErr& e) {
        if (d) d->err = e.msg;
        return RH_DN_ERR_HIP;
    } catch (const std::exception& e) {
        if (d) d->err = e.what();
        return RH_DN_ERR_HIP;
    }
}

}  // namespace

extern "C" {

int rh_dn_create(int64_t max_voxels, rh_denoiser** out) {
    if (!out || max_voxels <= 0) return RH_DN_ERR_ARG;
    *out = nullptr;
    rh_denoiser* d = new rh_denoiser();
    const int rc = guarded(d, [&] {
        DN_HIP(hipGetDevice(&d->device));
        DN_HIP(hipStreamCreate(&d->own_stream));
        d->stream = d->own_stream;
        d->static_scratch = static_cast<unsigned char*>(scratch_on(d->device));
        d->alloc(max_voxels);
        DN_HIP(hipStreamSynchronize(d->stream));
    });
    if (rc != RH_DN_OK) {
        fprintf(stderr, "rh_dn_create: %s\n", d->err.c_str());
        delete d;
        return rc;
    }
    *out = d;
    return RH_DN_OK;
}

void rh_dn_destroy(rh_denoiser* d) {
    if (!d) return;
    if (d->stream) (void)hipStreamSynchronize(d->stream);
    if (d->own_stream) (void)hipStreamSynchronize(d->own_stream);
    for (auto& pe : d->pending) {
        (void)hipEventDestroy(pe.second.first);
        (void)hipEventDestroy(pe.second.second);
    }
    hipStream_t s = d->own_stream;
    delete d;
    if (s) (void)hipStreamDestroy(s);
}

const char* rh_dn_last_error(const rh_denoiser* d) { return d ? d->err.c_str() : "null denoiser"; }

int rh_dn_set_weights(rh_denoiser* d, const float* w, int64_t n) {
    if (!d) return RH_DN_ERR_ARG;
    return guarded(d, [&] {
        if (!w || n != RH_DN_N_WEIGHTS) throw ArgErr{"expected 64800 weights"};
        DN_HIP(hipMemcpyAsync(d->w.p, w, n * sizeof(float), hipMemcpyHostToDevice, d->stream));
        d->reset_adam();
        d->pack();
        DN_HIP(hipStreamSynchronize(d->stream));
    });
}

int rh_dn_get_weights(rh_denoiser* d, float* w, int64_t n) {
    if (!d) return RH_DN_ERR_ARG;
    return guarded(d, [&] {
        if (!w || n != RH_DN_N_WEIGHTS) throw ArgErr{"expected 64800 weights"};
        DN_HIP(hipMemcpyAsync(w, d->w.p, n * sizeof(float), hipMemcpyDeviceToHost, d->stream));
        DN_HIP(hipStreamSynchronize(d->stream));
    });
}

int rh_dn_set_scales(rh_denoiser* d, const float scale[7]) {
    if (!d) return RH_DN_ERR_ARG;
    return guarded(d, [&] {
        if (!scale) throw ArgErr{"null scales"};
        for (int c = 0; c < 7; ++c)
            if (!(scale[c] > 0.0f) || !std::isfinite(scale[c])) throw ArgErr{"scales must be positive and finite"};
        memcpy(d->scale, scale, sizeof(d->scale));
    });
}

int rh_dn_forward(rh_denoiser* d, const float* x, float* y, int P, int X, int Y, int Z) {
    if (!d) return RH_DN_ERR_ARG;
    return guarded(d, [&] {
        if (!x || !y) throw ArgErr{"null volume"};
        if ((uintptr_t)y % 16) throw ArgErr{"the output volume must be 16-byte aligned"};
        d->check_dims(P, X, Y, Z);
        d->forward(x, y, P, X, Y, Z);
        if (d->timing) d->collect();
    });
}

int rh_dn_forward_static(rh_denoiser* d, const float* x, float* y, int P, int X, int Y, int Z) {
    if (!d) return RH_DN_ERR_ARG;
    return guarded(d, [&] {
        if (!x || !y) throw ArgErr{"null volume"};
        if ((uintptr_t)y % 16) throw ArgErr{"the output volume must be 16-byte aligned"};
        rh_denoiser::check_shape(P, X, Y, Z);
        if ((int64_t)P * X * Y * Z > RH_DN_STATIC_MAX_VOXELS)
            throw CapErr{"P*X*Y*Z = " + std::to_string((int64_t)P * X * Y * Z) +
                         " exceeds RH_DN_STATIC_MAX_VOXELS = " + std::to_string((int)RH_DN_STATIC_MAX_VOXELS) +
                         ", the size of the static scratch"};
        // the kernels run on the denoiser's stream, which must be on the
        // device whose copy of the scratch was looked up at creation
        int sdev = -1;
        DN_HIP(hipStreamGetDevice(d->stream, &sdev));
        if (sdev != d->device) throw ArgErr{"the denoiser's stream is on another device than the one it was created on"};
        bf16* x8s = reinterpret_cast<bf16*>(d->static_scratch);
        bf16* a1s = reinterpret_cast<bf16*>(d->static_scratch + STATIC_X8_BYTES);
        bf16* a2s = reinterpret_cast<bf16*>(d->static_scratch + STATIC_X8_BYTES + STATIC_ACT_BYTES);
        d->no_events = true;
        try {
            d->forward_with(x, y, P, X, Y, Z, x8s, a1s, a2s, a1s);
        } catch (...) {
            d->no_events = false;
            throw;
        }
        d->no_events = false;
    });
}

int rh_dn_gradient(rh_denoiser* d, const float* x, const float* target, int P, int X, int Y, int Z,
                   float* grad_host, float* loss_host) {
    if (!d) return RH_DN_ERR_ARG;
    return guarded(d, [&] {
        if (!x || !target || !grad_host) throw ArgErr{"null pointer"};
        d->check_dims(P, X, Y, Z);
        d->gradient(x, target, P, X, Y, Z);
        DN_HIP(hipMemcpyAsync(grad_host, d->g.p, RH_DN_N_WEIGHTS * sizeof(float), hipMemcpyDeviceToHost, d->stream));
        if (loss_host) DN_HIP(hipMemcpyAsync(loss_host, d->loss.p, sizeof(float), hipMemcpyDeviceToHost, d->stream));
        DN_HIP(hipStreamSynchronize(d->stream));
        if (d->timing) d->collect();
    });
}

int rh_dn_train_step(rh_denoiser* d, const float* x, const float* target, int P, int X, int Y, int Z, float lr,
                     float* loss_host) {
    if (!d) return RH_DN_ERR_ARG;
    return guarded(d, [&] {
        if (!x || !target) throw ArgErr{"null pointer"};
        if (!(lr >= 0.0f) || !std::isfinite(lr)) throw ArgErr{"learning rate must be finite and >= 0"};
        d->check_dims(P, X, Y, Z);
        d->gradient(x, target, P, X, Y, Z);
        d->adam(lr);
        if (loss_host) {
            DN_HIP(hipMemcpyAsync(loss_host, d->loss.p, sizeof(float), hipMemcpyDeviceToHost, d->stream));
            DN_HIP(hipStreamSynchronize(d->stream));
        }
        if (d->timing) d->collect();
    });
}

int rh_dn_synchronize(rh_denoiser* d) {
    if (!d) return RH_DN_ERR_ARG;
    return guarded(d, [&] {
        DN_HIP(hipStreamSynchronize(d->stream));
        if (d->timing) d->collect();
    });
}

int rh_dn_set_stream(rh_denoiser* d, void* hip_stream) {
    if (!d) return RH_DN_ERR_ARG;
    return guarded(d, [&] {
        hipStream_t ns = hip_stream ? (hipStream_t)hip_stream : d->own_stream;
        if (ns == d->stream) return;
        // the work already queued finishes before later work on the new stream starts
        DN_HIP(hipStreamSynchronize(d->stream));
        d->stream = ns;
    });
}

int rh_dn_set_timing(rh_denoiser* d, int on) {
    if (!d) return RH_DN_ERR_ARG;
    return guarded(d, [&] {
        d->collect();
        d->timing = on != 0;
    });
}

int rh_dn_kernel_time(rh_denoiser* d, const char* name, double* ms, int64_t* launches) {
    if (!d || !name || !ms || !launches) return RH_DN_ERR_ARG;
    return guarded(d, [&] {
        d->collect();
        auto it = d->times.find(name);
        *ms = it == d->times.end() ? 0.0 : it->second.first;
        *launches = it == d->times.end() ? 0 : it->second.second;
    });
}

int rh_dn_reset_timing(rh_denoiser* d) {
    if (!d) return RH_DN_ERR_ARG;
    return guarded(d, [&] {
        d->collect();
        d->times.clear();
    });
}

}  // extern "C"
