// rh_capi.hip -- the host side: a simulation context per domain, the
// operators of the reference's step as sequences of kernel launches, and the
// extern "C" entry points declared in include/rh_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rh_gpu.h"
#include "../../include/rh_denoise.h"
#include "rh_fields.hpp"
#include "rh_geom.hpp"
#include "rh_inject.hpp"
#include "rh_log.hpp"
#include "rh_vlsv.hpp"
#include "rh_particles.hpp"

namespace rh {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct CapError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct StateError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define RH_HIP(call)                                                                               \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw HipError(std::string(#call) + ": " + hipGetErrorString(e_));                     \
    } while (0)

static const int FIELD_DIM[RH_N_FIELDS] = {3, 3, 3, 3, 3, 3, 1, 1, 3, 3, 3, 3, 3, 1};

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { free(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    void alloc(size_t count) {
        free();
        if (count == 0) count = 1;
        RH_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { free(); }
};

struct Timer {
    struct Pending {
        hipEvent_t e0, e1;
        int species;  // the species the launch worked on, -1 for none
    };
    std::vector<Pending> pending;
    double ms = 0;
    int64_t launches = 0;
};

class Ctx {
public:
    virtual ~Ctx() = default;
    std::string err;
    virtual int add_species(const rh_species_desc* d) = 0;
    virtual void upload_field(int id, const double* data) = 0;
    virtual void download_field(int id, double* data) = 0;
    virtual void upload_flags(int id, const uint8_t* data) = 0;
    virtual void upload_particles(int s, int64_t n, const int64_t* cell, const double* st) = 0;
    virtual int64_t num_particles(int s) = 0;
    virtual void download_particles(int s, int64_t cap, int64_t* cell, double* st) = 0;
    virtual void propagate(int s) = 0;
    virtual void wait_particle_sends(int s) = 0;
    virtual void clear_accumulation() = 0;
    virtual void accumulate(int s) = 0;
    virtual void add_remote_updates() = 0;
    virtual void apply_bcs(int s) = 0;
    virtual void propagate_B() = 0;
    virtual void step(int n) = 0;
    virtual void diagnostics(int s, double* out) = 0;
    virtual void counters(double* out) = 0;
    virtual void denoise(rh_denoiser* dn, double max_rel) = 0;
    virtual void set_denoiser(rh_denoiser* dn, double max_rel) = 0;
    virtual void synchronize() = 0;
    virtual void set_timing(bool on) = 0;
    virtual void set_timing_only(const char* name) = 0;
    virtual void set_species_streams(bool on) = 0;
    virtual void kernel_time(const char* name, double* ms, int64_t* launches) = 0;
    virtual void reset_timing() = 0;
    virtual void layout(int64_t* out) = 0;
    virtual void set_injector(int s, const rh_injector_desc* d) = 0;
    virtual void inject(int s) = 0;
    virtual void set_timestep(int64_t t) = 0;
    virtual int64_t timestep() = 0;
    virtual int64_t injected(int s) = 0;
    virtual void log_begin(double t) = 0;
    virtual void log_values(double t, int include_inner, double* field, double* pop) = 0;
    virtual void log_open(const char* dir, const char* const* names, int precision, int include_inner,
                          double max_B_stop) = 0;
    virtual void log_write(double t) = 0;
    virtual void write_vlsv(const char* path, double t, int64_t timestep, const char* const* vars, int nvars,
                            const char* const* pops) = 0;
    virtual void set_temporal_averages(bool on) = 0;
};

// Boltzmann constant of vth = sqrt(kB T / m) (CODATA 2002, the same table as
// the e and m_p of tools/trace_test_particles.py:47-48).
constexpr double KB = 1.3806505e-23;

// Constants of sampleDerivGaussian (src/particle_injector.cpp:97-106) for x0.
static void deriv_gauss_consts(double x0, double& xm, double& c, double& d, double& cxm) {
#pragma clang fp contract(off)
    xm = 0.5 * (x0 + std::sqrt(x0 * x0 + 4.0));
    c = 1.0 / (std::exp(-0.5 * x0 * x0) + x0 * std::sqrt(0.5 * M_PI) * std::erfc(-x0 * (1.0 / std::sqrt(2.0))));
    d = (x0 - xm) * (x0 - xm);
    cxm = c * xm;
}

// Moments <-> denoiser volume [Z][Y][X][7] (n, Ji, cellB) and back; the
// cell index n = i + nx*(j + ny*k) is the voxel index of the volume.  The
// gather also writes the binomially filtered charge density (weights 1/4,
// 1/2, 1/4 along each axis; a missing neighbour's weight goes to the cell
// itself), the reference the fallback bound is measured against.
template <typename Real>
__global__ void k_dn_gather(Geom g, const Real* __restrict__ rho, V3<Real> ji, V3<Real> b, float* __restrict__ out,
                            float* __restrict__ filt) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nc) return;
    float* o = out + (size_t)n * 7;
    o[0] = (float)rho[n];
    o[1] = (float)ji.x[n]; o[2] = (float)ji.y[n]; o[3] = (float)ji.z[n];
    o[4] = (float)b.x[n]; o[5] = (float)b.y[n]; o[6] = (float)b.z[n];
    int i, j, k;
    cell_ijk(g, n, i, j, k);
    const float w1[3] = {0.25f, 0.5f, 0.25f};
    float acc = 0.0f;
    for (int c = -1; c <= 1; ++c)
        for (int bb = -1; bb <= 1; ++bb)
            for (int a = -1; a <= 1; ++a) {
                const int m = nbr(g, i, j, k, a, bb, c);
                acc += w1[a + 1] * w1[bb + 1] * w1[c + 1] * (float)rho[m >= 0 ? m : n];
            }
    filt[n] = acc;
}

// Fallback bound: a cell whose denoised charge density differs from the
// binomially filtered deposit by more than max_rel * |filtered| keeps its
// deposited moments (also for a non-finite output) and is counted.
template <typename Real>
__global__ void k_dn_scatter(int nc, const float* __restrict__ y, const float* __restrict__ filt, double max_rel,
                             double min_rho, Real* __restrict__ rho, V3<Real> ji,
                             unsigned long long* __restrict__ counters) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    bool fallback = false;
    if (n < nc) {
        const float* q = y + (size_t)n * 4;
        const double f = (double)filt[n], r1 = (double)q[0];
        if (!(fabs(r1 - f) <= max_rel * fabs(f))) {
            fallback = true;
        } else {
            rho[n] = (Real)(r1 < min_rho ? min_rho : r1);
            ji.x[n] = (Real)q[1]; ji.y[n] = (Real)q[2]; ji.z[n] = (Real)q[3];
        }
    }
    count_event(counters, C_DN_FALLBACK, fallback);
}

template <typename Real>
class Sim final : public Ctx {
public:
    // Per species: the sorted store (a, sub) and the destination of the next
    // sort (d, sub_new), swapped after every sort; cnt_tile holds the 16-bit
    // slot counts of pass 0 and then, for pass 1, each slot's offset inside
    // its destination subcell's range (sub_new holds the ranges' starts).
    struct Species {
        rh_species_desc d;
        size_t cap = 0;
        DevBuf<Real> a[6], dst[6];
        DevBuf<uint8_t> a_lc, dst_lc;
        DevBuf<uint32_t> sub, sub_new;
        DevBuf<uint16_t> cnt_tile;
        DevBuf<Real> partial;  // deposit: per-tile ghost partials (ntiles x next x 4)
        // per tile, 1 if it held particles at the push (its slot table was
        // written)
        DevBuf<uint8_t> live_src;
        DevBuf<double> ave_n, ave_v;  // temporal averages: running charge and current sums (FP64)
        bool deposited = false;
        bool pushed = false;  // pass 0 done, pass 1 pending
        PushParams pp{};      // parameters of the pending push
        // injector: listed cells in subcell-key order, their rates
        // (exosphere), per tile the first listed cell, per listed cell the
        // offset of its new particles past the store's end
        InjParams inj{};
        int inj_ncell = 0;
        int inj_slots = 0;  // solar wind: fixed places per listed cell (ceil of the rate)
        DevBuf<int> inj_cells;
        DevBuf<double> inj_rate;
        DevBuf<uint32_t> inj_cbeg, inj_off;
        DevBuf<unsigned long long> inj_total;
        bool inj_pending = false;  // injected particles wait past the store's end
        // at least 4 particles per cell at the end of the last rh_step call
        // (the move's occupancy cap pays for such species only)
        bool dense = true;
        // log counters: the push's own counters (maxVi per species), per
        // boundary cell (escapes, sum v^2, impacts, sum v^2), per injection
        // place sum v^2; host carries of the accumulators when their layout
        // changes; values at the last log line
        DevBuf<unsigned long long> pcount;
        DevBuf<double> bc_acc, inj_acc;
        double carry_bc[4] = {0, 0, 0, 0};
        double carry_inj_v2 = 0;
        double log_maxvi0 = 0;
        int64_t log_inj0 = 0;
        InjAppendix appendix() const {
            return inj_pending ? InjAppendix{inj_cbeg.p, inj_off.p} : InjAppendix{nullptr, nullptr};
        }
        PArrays<Real> A() { return {a[0].p, a[1].p, a[2].p, a[3].p, a[4].p, a[5].p, a_lc.p}; }
        PArrays<Real> Dst() { return {dst[0].p, dst[1].p, dst[2].p, dst[3].p, dst[4].p, dst[5].p, dst_lc.p}; }
        void swap_stores() {
            for (int q = 0; q < 6; ++q) std::swap(a[q].p, dst[q].p);
            std::swap(a_lc.p, dst_lc.p);
            std::swap(sub.p, sub_new.p);
        }
        // the species' own stream for the particle phase of a step, the event
        // that marks the end of its work there, and its scan's block sums
        hipStream_t ss = nullptr;
        hipEvent_t done = nullptr;
        std::vector<DevBuf<uint32_t>> scan_lv;
        Species() = default;
        Species(const Species&) = delete;
        Species& operator=(const Species&) = delete;
        ~Species() {
            if (ss) (void)hipStreamSynchronize(ss);
            if (done) (void)hipEventDestroy(done);
            if (ss) (void)hipStreamDestroy(ss);
        }
    };

    Geom g{};
    rh_params P{};
    hipStream_t stream = nullptr;
    hipStream_t ls = nullptr;      // the stream launches go to (stream, or a species' own)
    hipEvent_t fork_ev = nullptr;  // the particle phase's species streams wait for it
    bool species_streams = false;  // step(): the species' particle phases side by side
    DevBuf<Real> fld[RH_N_FIELDS];
    DevBuf<Real> tmpv;            // old nodeE for the gaussian smoothing
    DevBuf<Real> faceB_alt;       // the fused field update's second faceB buffer
    DevBuf<uint8_t> flg[RH_N_FLAGS];
    DevBuf<uint8_t> flg_fu;       // the fused field update's three inner flags in one byte per cell
    DevBuf<int> ext_cells, bc_cells;
    int n_ext = 0, n_bc = 0;
    DevBuf<unsigned long long> dcount;
    DevBuf<double> diag_partial;
    std::vector<std::unique_ptr<Species>> sp;
    std::vector<DevBuf<uint32_t>> scan_lv;  // per level: block sums
    DevBuf<uint32_t> zero_u32;
    DevBuf<uint16_t> dep_list;   // deposit gather lists of subcell terms (next x 64; crowded tiles)
    DevBuf<uint8_t> dep_count;
    DevBuf<uint16_t> dep_tlist;  // deposit gather lists of cell-target terms (next x 32)
    DevBuf<uint8_t> dep_tcount;
    int kbits = 12;
    bool timing = false;
    std::string timing_only;  // when not empty, only launches of this name are timed
    std::map<std::string, Timer> timers;
    std::vector<hipEvent_t> event_pool;
    std::vector<uint8_t> h_inner_particle;

    Sim(const rh_grid_desc* gd, const rh_params* pr) {
        if (gd->nx < 3 || gd->ny < 3 || gd->nz < 3) throw ArgError("each grid dimension needs >= 3 cells");
        if (!(gd->dx > 0)) throw ArgError("dx must be positive");
        int dev_count = 0;
        if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) throw HipError("no GPU visible");
        P = *pr;
        g.nx = gd->nx; g.ny = gd->ny; g.nz = gd->nz;
        g.nc = g.nx * g.ny * g.nz;
        for (int d = 0; d < 3; ++d) g.per[d] = gd->periodic[d] ? 1 : 0;
        g.dx = gd->dx; g.x0 = gd->x0; g.y0 = gd->y0; g.z0 = gd->z0;
        {
            int dev = 0, lds = 0;
            RH_HIP(hipGetDevice(&dev));
            RH_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
            lds_limit = (size_t)lds;
            int lds_cu = 0;
            RH_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
            lds_per_cu = (size_t)lds_cu;
        }
        choose_tiles(gd->tile);
        // dynamic LDS above the default 64 KiB (FP64 tiles of 64 cells)
        RH_HIP(hipFuncSetAttribute((const void*)k_deposit<Real>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)deposit_lds_bytes<Real>(g)));
        RH_HIP(hipFuncSetAttribute((const void*)k_push_move<Real>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)std::max(push_move_lds_bytes<Real>(g), move_lds_bytes())));
        RH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        ls = stream;
        RH_HIP(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
        for (int f = 0; f < RH_N_FIELDS; ++f) {
            fld[f].alloc((size_t)g.nc * FIELD_DIM[f]);
            fill(fld[f].p, Real(0), fld[f].n);
        }
        tmpv.alloc((size_t)g.nc * 3);
        faceB_alt.alloc((size_t)g.nc * 3);
        for (int f = 0; f < RH_N_FLAGS; ++f) {
            flg[f].alloc(g.nc);
            RH_HIP(hipMemsetAsync(flg[f].p, 0, g.nc, stream));
        }
        flg_fu.alloc(g.nc);
        RH_HIP(hipMemsetAsync(flg_fu.p, 0, g.nc, stream));
        h_inner_particle.assign(g.nc, 0);
        // exterior cells: missing any neighbour
        std::vector<int> ext;
        for (int n = 0; n < g.nc; ++n) {
            int i, j, k;
            cell_ijk(g, n, i, j, k);
            if (!all_nbrs(g, i, j, k)) ext.push_back(n);
        }
        n_ext = (int)ext.size();
        ext_cells.alloc(std::max<size_t>(ext.size(), 1));
        if (n_ext) RH_HIP(hipMemcpy(ext_cells.p, ext.data(), ext.size() * sizeof(int), hipMemcpyHostToDevice));
        rebuild_bc_cells();
        dcount.alloc((size_t)C_N * COUNTER_SLOTS);
        RH_HIP(hipMemsetAsync(dcount.p, 0, sizeof(unsigned long long) * C_N * COUNTER_SLOTS, stream));
        diag_partial.alloc(DIAG_BLOCKS * DIAG_Q);
        {
            std::vector<uint16_t> hl((size_t)g.next * 64, 0);
            std::vector<uint8_t> hc((size_t)g.next, 0);
            deposit_gather_lists(g, hl.data(), hc.data());
            dep_list.alloc(hl.size());
            dep_count.alloc(hc.size());
            RH_HIP(hipMemcpy(dep_list.p, hl.data(), hl.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
            RH_HIP(hipMemcpy(dep_count.p, hc.data(), hc.size(), hipMemcpyHostToDevice));
            std::vector<uint16_t> tl((size_t)g.next * 32, 0);
            std::vector<uint8_t> tc((size_t)g.next, 0);
            deposit_target_lists(g, tl.data(), tc.data());
            dep_tlist.alloc(tl.size());
            dep_tcount.alloc(tc.size());
            RH_HIP(hipMemcpy(dep_tlist.p, tl.data(), tl.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
            RH_HIP(hipMemcpy(dep_tcount.p, tc.data(), tc.size(), hipMemcpyHostToDevice));
        }
        zero_u32.alloc(1);
        RH_HIP(hipMemsetAsync(zero_u32.p, 0, sizeof(uint32_t), stream));
        alloc_scan_levels(scan_lv);
        kbits = 1;
        while ((1 << kbits) < g.extk) ++kbits;
        RH_HIP(hipStreamSynchronize(stream));
    }

    // block sums of a scan over nkeys elements, per level
    void alloc_scan_levels(std::vector<DevBuf<uint32_t>>& lv) {
        size_t n = (size_t)g.nkeys;
        while (n > (size_t)SCAN_TILE) {
            const size_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
            lv.emplace_back();
            lv.back().alloc(nb + 1);
            lv.emplace_back();
            lv.back().alloc(nb + 1);
            n = nb;
        }
    }

    ~Sim() override {
        for (auto& s_ : sp)
            if (s_->ss) (void)hipStreamSynchronize(s_->ss);
        if (stream) (void)hipStreamSynchronize(stream);
        if (dn_attached) (void)rh_dn_set_stream(dn_attached, nullptr);  // before the stream goes
        for (auto& t : timers)
            for (auto& pe : t.second.pending) { (void)hipEventDestroy(pe.e0); (void)hipEventDestroy(pe.e1); }
        for (auto e : event_pool) (void)hipEventDestroy(e);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        sp.clear();
        if (stream) (void)hipStreamDestroy(stream);
    }

    size_t lds_limit = 65536;  // the device's LDS per workgroup (queried at creation)
    size_t lds_per_cu = 0;     // and per CU

    // The move pass (pass 1 of k_push_move) runs with at most MOVE_WG_PER_CU
    // workgroups per CU, set through its dynamic LDS request: with fewer
    // tiles' scattered stores in flight per CU it runs faster (the 128x96x96
    // box: 500 us per launch at the 7 workgroups its registers allow, 483 at
    // 5 or 6, 474 at 4, 526 at 3).  The push pass keeps the full occupancy,
    // and so does the move of a species with fewer than 4 particles per cell
    // (at the end of the last rh_step call): its move is set by a few crowded
    // tiles, which the cap only starts later.
    static constexpr int MOVE_WG_PER_CU = 4;
    size_t move_lds_bytes(const Species* S_ = nullptr) const {
        const size_t need = push_move_lds_bytes<Real>(g);
        if (lds_per_cu == 0 || (S_ && !S_->dense)) return need;
        const size_t cap = lds_per_cu / MOVE_WG_PER_CU - 1024;  // below the share, so MOVE_WG_PER_CU fit
        return std::min(lds_limit, std::max(need, cap));
    }

    // LDS of the largest per-tile kernels for a tile shape
    size_t tile_lds_bytes(const int* t) const {
        Geom h = g;
        h.tx = t[0]; h.ty = t[1]; h.tz = t[2];
        h.tc = t[0] * t[1] * t[2];
        h.ex = t[0] + 2; h.ey = t[1] + 2; h.ez = t[2] + 2;
        h.next = h.ex * h.ey * h.ez;
        h.extk = h.next * 8;
        return std::max(deposit_lds_bytes<Real>(h), push_move_lds_bytes<Real>(h));
    }

    void choose_tiles(const int32_t* req) {
        int t[3] = {req ? req[0] : 0, req ? req[1] : 0, req ? req[2] : 0};
        const int dims[3] = {g.nx, g.ny, g.nz};
        // the deposit runs one thread per subcell (at most 512) and both
        // per-tile kernels keep their tile in LDS
        const int max_cells = 64;
        if (t[0] <= 0 || t[1] <= 0 || t[2] <= 0) {
            for (int d = 0; d < 3; ++d) {
                t[d] = 4;
                while (t[d] > 1 && dims[d] % t[d]) t[d] /= 2;
            }
            auto ntiles = [&]() { return (dims[0] / t[0]) * (dims[1] / t[1]) * (dims[2] / t[2]); };
            while (t[0] * t[1] * t[2] > max_cells || (t[0] * t[1] * t[2] > 1 && tile_lds_bytes(t) > lds_limit)) {
                int d = 0;
                for (int q = 1; q < 3; ++q) if (t[q] > t[d]) d = q;
                t[d] /= 2;
            }
            // enough tiles to fill the chip when the grid is small
            while (ntiles() < 4096) {
                int d = -1;
                for (int q = 0; q < 3; ++q)
                    if (t[q] > 2 && (d < 0 || t[q] > t[d])) d = q;
                if (d < 0) break;
                t[d] /= 2;
            }
        }
        for (int d = 0; d < 3; ++d)
            if (t[d] < 1 || (t[d] & (t[d] - 1)) || dims[d] % t[d])
                throw ArgError("tile sides must be powers of two that divide the grid");
        if (t[0] * t[1] * t[2] > max_cells) throw ArgError("tile has more than 64 cells");
        if (tile_lds_bytes(t) > lds_limit)
            throw ArgError("tile needs " + std::to_string(tile_lds_bytes(t)) + " bytes of LDS, the device has " +
                           std::to_string(lds_limit));
        g.tx = t[0]; g.ty = t[1]; g.tz = t[2];
        g.ltx = __builtin_ctz(g.tx); g.lty = __builtin_ctz(g.ty); g.ltz = __builtin_ctz(g.tz);
        g.tc = g.tx * g.ty * g.tz;
        g.ntx = g.nx / g.tx; g.nty = g.ny / g.ty; g.ntz = g.nz / g.tz;
        g.ntiles = g.ntx * g.nty * g.ntz;
        xcd_bricks(g);
        g.ex = g.tx + 2; g.ey = g.ty + 2; g.ez = g.tz + 2;
        g.next = g.ex * g.ey * g.ez;
        g.extk = g.next * 8;
        if (g.extk > MAX_EXTK) throw ArgError("tile extended region too large");
        const uint64_t nk = (uint64_t)g.ntiles * g.tc * 8;
        if (nk >= 0xFFFFFFFFull) throw ArgError("grid too large for 32-bit subcell keys");
        if ((uint64_t)g.ntiles * g.extk >= 0xFFFFFFFFull) throw ArgError("tiles too small for this grid (slot table)");
        g.nkeys = (uint32_t)nk;
    }

    void rebuild_bc_cells() {
        std::vector<int> cells;
        for (int n = 0; n < g.nc; ++n) {
            int i, j, k;
            cell_ijk(g, n, i, j, k);
            if (!all_nbrs(g, i, j, k) || h_inner_particle[n]) cells.push_back(n);
        }
        n_bc = (int)cells.size();
        bc_cells.alloc(std::max<size_t>(cells.size(), 1));
        if (n_bc) RH_HIP(hipMemcpy(bc_cells.p, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice));
    }

    // ---- launch helpers -------------------------------------------------

    hipEvent_t get_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e;
        RH_HIP(hipEventCreate(&e));
        return e;
    }

    // A launch made on behalf of one species is also counted under
    // "<name>:<species>" (e.g. "move:1"), so that per-species times can be
    // read where species differ in size.
    void resolve_timers() {
        RH_HIP(hipStreamSynchronize(stream));
        std::map<std::string, Timer> per_species;
        for (auto& kv : timers) {
            for (auto& pe : kv.second.pending) {
                float ms = 0;
                RH_HIP(hipEventElapsedTime(&ms, pe.e0, pe.e1));
                kv.second.ms += ms;
                if (pe.species >= 0) {
                    Timer& t = per_species[kv.first + ":" + std::to_string(pe.species)];
                    t.ms += ms;
                    t.launches++;
                }
                event_pool.push_back(pe.e0);
                event_pool.push_back(pe.e1);
            }
            kv.second.pending.clear();
        }
        for (auto& kv : per_species) {
            Timer& t = timers[kv.first];
            t.ms += kv.second.ms;
            t.launches += kv.second.launches;
        }
    }

    int timing_species = -1;  // the species the operator being run works on
    struct SpeciesScope {
        int& slot;
        int prev;
        SpeciesScope(int& s, int species) : slot(s), prev(s) { slot = species; }
        ~SpeciesScope() { slot = prev; }
    };

    template <typename K, typename... Args>
    void launch(const char* name, K kernel, dim3 grid, dim3 block, size_t shmem, Args... args) {
        if ((size_t)grid.x * grid.y * grid.z == 0) return;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        const bool timed = timing && (timing_only.empty() || timing_only == name);
        if (timed) {
            e0 = get_event();
            e1 = get_event();
            RH_HIP(hipEventRecord(e0, ls));
        }
        hipLaunchKernelGGL(kernel, grid, block, shmem, ls, args...);
        RH_HIP(hipGetLastError());
        if (timed) {
            RH_HIP(hipEventRecord(e1, ls));
            Timer& t = timers[name];
            t.pending.push_back({e0, e1, timing_species});
            t.launches++;
            if (t.pending.size() > 4096) resolve_timers();
        }
    }

    template <typename T>
    void fill(T* p, T v, size_t n) {
        if (n == 0) return;
        const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
        launch("fill", k_fill<T>, blocks, 256, 0, p, v, n);
    }

    static unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }
    // per-tile launches: one workgroup per tile in XCD-aware order (xcd_tile)
    dim3 tile_grid() const { return dim3(N_XCD * g.xchunk); }
    dim3 cell_grid() const { return dim3(cdiv(g.nx, 32), cdiv(g.ny, 8), g.nz); }
    static dim3 cell_block() { return dim3(32, 8, 1); }

    V3<Real> v3(int id) {
        Real* p = fld[id].p;
        return {p, p + g.nc, p + 2 * (size_t)g.nc};
    }

    Species& S(int s) {
        if (s < 0 || s >= (int)sp.size()) throw ArgError("bad species id");
        return *sp[s];
    }

    // ---- set-up and data movement ----------------------------------------

    int add_species(const rh_species_desc* d) override {
        auto s = std::make_unique<Species>();
        s->d = *d;
        s->sub.alloc((size_t)g.nkeys + 1);
        s->sub_new.alloc((size_t)g.nkeys + 1);
        RH_HIP(hipMemsetAsync(s->sub.p, 0, sizeof(uint32_t) * ((size_t)g.nkeys + 1), stream));
        s->cnt_tile.alloc((size_t)g.ntiles * g.extk);
        s->partial.alloc((size_t)g.ntiles * g.next * 4);
        s->live_src.alloc((size_t)g.ntiles);
        fill(s->live_src.p, (uint8_t)0, s->live_src.n);
        s->pcount.alloc((size_t)C_N * COUNTER_SLOTS);
        RH_HIP(hipMemsetAsync(s->pcount.p, 0, sizeof(unsigned long long) * C_N * COUNTER_SLOTS, stream));
        s->bc_acc.alloc((size_t)std::max(n_bc, 1) * 4);
        fill(s->bc_acc.p, 0.0, s->bc_acc.n);
        ensure_capacity(*s, (size_t)std::max<int64_t>(d->capacity, 0));
        ensure_averages(*s);
        alloc_scan_levels(s->scan_lv);
        sp.push_back(std::move(s));
        RH_HIP(hipStreamSynchronize(stream));
        return (int)sp.size() - 1;
    }

    void ensure_capacity(Species& s, size_t cap) {
        if (cap <= s.cap && s.cap > 0) return;
        if (cap >= 0xFFFFFFFFull) throw CapError("particle capacity exceeds 32-bit indices");
        cap = std::max<size_t>(cap, 64);
        for (int q = 0; q < 6; ++q) { s.a[q].alloc(cap); s.dst[q].alloc(cap); }
        s.a_lc.alloc(cap);
        s.dst_lc.alloc(cap);
        s.cap = cap;
    }

    void upload_field(int id, const double* data) override {
        if (id < 0 || id >= RH_N_FIELDS || !data) throw ArgError("bad field id or null data");
        const int dim = FIELD_DIM[id];
        std::vector<Real> h((size_t)g.nc * dim);
        for (int l = 0; l < dim; ++l)
            for (int n = 0; n < g.nc; ++n) h[(size_t)l * g.nc + n] = (Real)data[(size_t)n * dim + l];
        RH_HIP(hipMemcpyAsync(fld[id].p, h.data(), h.size() * sizeof(Real), hipMemcpyHostToDevice, stream));
        RH_HIP(hipStreamSynchronize(stream));
        pf_dirty = true;
    }

    void download_field(int id, double* data) override {
        if (id < 0 || id >= RH_N_FIELDS || !data) throw ArgError("bad field id or null data");
        const int dim = FIELD_DIM[id];
        std::vector<Real> h((size_t)g.nc * dim);
        RH_HIP(hipMemcpyAsync(h.data(), fld[id].p, h.size() * sizeof(Real), hipMemcpyDeviceToHost, stream));
        RH_HIP(hipStreamSynchronize(stream));
        for (int l = 0; l < dim; ++l)
            for (int n = 0; n < g.nc; ++n) data[(size_t)n * dim + l] = (double)h[(size_t)l * g.nc + n];
    }

    void upload_flags(int id, const uint8_t* data) override {
        if (id < 0 || id >= RH_N_FLAGS || !data) throw ArgError("bad flag id or null data");
        RH_HIP(hipMemcpyAsync(flg[id].p, data, g.nc, hipMemcpyHostToDevice, stream));
        if (id == RH_INNER_FIELD || id == RH_INNER_NODE || id == RH_INNER_CELLEP)
            launch("fill", k_pack_inner_flags, cdiv(g.nc, 256), 256, 0, g.nc, (const uint8_t*)flg[RH_INNER_FIELD].p,
                   (const uint8_t*)flg[RH_INNER_NODE].p, (const uint8_t*)flg[RH_INNER_CELLEP].p, flg_fu.p);
        RH_HIP(hipStreamSynchronize(stream));
        if (id == RH_INNER_PARTICLE) {
            for (int n = 0; n < g.nc; ++n) h_inner_particle[n] = data[n] ? 1 : 0;
            for (auto& sp_ : sp) carry_acc(*sp_, true, false);
            rebuild_bc_cells();
            for (auto& sp_ : sp) {
                sp_->bc_acc.alloc((size_t)std::max(n_bc, 1) * 4);
                fill(sp_->bc_acc.p, 0.0, sp_->bc_acc.n);
            }
            RH_HIP(hipStreamSynchronize(stream));
        }
    }

    static Real to_unit(double x, double dx) {
        Real r = (Real)(x / dx);
        if (r >= Real(1)) r = std::nextafter(Real(1), Real(0));
        if (r < Real(0)) r = Real(0);
        return r;
    }

    void upload_particles(int s, int64_t n, const int64_t* cell, const double* st) override {
        Species& S_ = S(s);
        if (n < 0 || (n > 0 && (!cell || !st))) throw ArgError("bad particle arrays");
        if ((uint64_t)n >= 0xFFFFFFFFull) throw CapError("too many particles for 32-bit indices");
        // keys and a stable counting sort on the host
        std::vector<uint32_t> key((size_t)n);
        std::vector<Real> u[3];
        for (int q = 0; q < 3; ++q) u[q].resize((size_t)n);
        for (int64_t p = 0; p < n; ++p) {
            if (cell[p] < 0 || cell[p] >= g.nc) throw ArgError("particle cell out of range");
            int i, j, k;
            cell_ijk(g, (int)cell[p], i, j, k);
            const Real X = to_unit(st[7 * p + 0], g.dx), Y = to_unit(st[7 * p + 1], g.dx), Z = to_unit(st[7 * p + 2], g.dx);
            u[0][p] = X; u[1][p] = Y; u[2][p] = Z;
            const int o = (X >= Real(0.5)) + 2 * (Y >= Real(0.5)) + 4 * (Z >= Real(0.5));
            key[p] = cell_key(g, i, j, k) * 8u + (uint32_t)o;
        }
        std::vector<uint32_t> sub((size_t)g.nkeys + 1, 0);
        for (int64_t p = 0; p < n; ++p) sub[key[p] + 1]++;
        for (size_t k2 = 0; k2 < g.nkeys; ++k2) sub[k2 + 1] += sub[k2];
        std::vector<uint32_t> fillp(sub.begin(), sub.end() - 1);
        std::vector<Real> h[6];
        for (int q = 0; q < 6; ++q) h[q].resize((size_t)n);
        std::vector<uint8_t> hl((size_t)n);
        for (int64_t p = 0; p < n; ++p) {
            const uint32_t d = fillp[key[p]]++;
            hl[d] = (uint8_t)((key[p] >> 3) % (uint32_t)g.tc);
            h[0][d] = u[0][p]; h[1][d] = u[1][p]; h[2][d] = u[2][p];
            h[3][d] = (Real)st[7 * p + 3]; h[4][d] = (Real)st[7 * p + 4]; h[5][d] = (Real)st[7 * p + 5];
        }
        size_t cap = (size_t)std::max<int64_t>(S_.d.capacity, 0);
        if (cap < (size_t)n) cap = (size_t)n + (size_t)n / 2;
        ensure_capacity(S_, cap);
        for (int q = 0; q < 6; ++q)
            RH_HIP(hipMemcpyAsync(S_.a[q].p, h[q].data(), sizeof(Real) * n, hipMemcpyHostToDevice, stream));
        RH_HIP(hipMemcpyAsync(S_.a_lc.p, hl.data(), n, hipMemcpyHostToDevice, stream));
        RH_HIP(hipMemcpyAsync(S_.sub.p, sub.data(), sizeof(uint32_t) * sub.size(), hipMemcpyHostToDevice, stream));
        RH_HIP(hipStreamSynchronize(stream));
        S_.pushed = false;
        S_.inj_pending = false;
    }

    // live particles (those removed by the boundary conditions stay stored,
    // marked, until the next sort; injected ones waiting for their first
    // push are included)
    int64_t num_particles(int s) override {
        double t[DIAG_Q];
        diag_raw(s, t);
        return (int64_t)t[0];
    }

    void download_particles(int s, int64_t capn, int64_t* cell, double* st) override {
        Species& S_ = S(s);
        if (S_.pushed) throw StateError("download between propagate and wait_particle_sends");
        flush_injected(S_);
        std::vector<uint32_t> sub((size_t)g.nkeys + 1);
        RH_HIP(hipMemcpyAsync(sub.data(), S_.sub.p, sizeof(uint32_t) * sub.size(), hipMemcpyDeviceToHost, stream));
        RH_HIP(hipStreamSynchronize(stream));
        const uint32_t np = sub[g.nkeys];
        std::vector<Real> h[6];
        for (int q = 0; q < 6; ++q) {
            h[q].resize(np);
            if (np) RH_HIP(hipMemcpy(h[q].data(), S_.a[q].p, sizeof(Real) * np, hipMemcpyDeviceToHost));
        }
        // the cell of a particle follows from the subcell range holding it
        std::vector<uint32_t> hc(np);
        for (uint32_t key = 0; key < g.nkeys; ++key) {
            if (sub[key] == sub[key + 1]) continue;
            const int T = (int)(key / (uint32_t)(g.tc * 8)), lc = (int)((key % (uint32_t)(g.tc * 8)) >> 3);
            int ox, oy, oz;
            tile_origin(g, T, ox, oy, oz);
            const uint32_t c = (uint32_t)cell_index(g, ox + lc % g.tx, oy + (lc / g.tx) % g.ty, oz + lc / (g.tx * g.ty));
            for (uint32_t p = sub[key]; p < sub[key + 1]; ++p) hc[p] = c;
        }
        int64_t out = 0;
        for (uint32_t p = 0; p < np; ++p) {
            if (h[0][p] < Real(0)) continue;
            if (out >= capn) throw CapError("download buffer too small");
            if (cell) cell[out] = hc[p];
            if (st) {
                double* q = st + 7 * out;
                q[0] = (double)h[0][p] * g.dx; q[1] = (double)h[1][p] * g.dx; q[2] = (double)h[2][p] * g.dx;
                q[3] = (double)h[3][p]; q[4] = (double)h[4][p]; q[5] = (double)h[5][p];
                q[6] = S_.d.w;
            }
            ++out;
        }
    }

    // ---- the operators ------------------------------------------------------

    void propagate(int s) override {
        SpeciesScope species_scope(timing_species, s);
        Species& S_ = S(s);
        if (S_.pushed) throw StateError("propagate called twice without wait_particle_sends");
        PushParams pp;
        pp.dt = P.dt;
        pp.dt_dx = P.dt / g.dx;
        pp.qmdt2 = 0.5 * S_.d.q * P.dt / S_.d.m;
        pp.max_vi2 = P.max_ion_speed2;
        pp.GMdt = P.GMdt;
        pp.accelerate = S_.d.accelerate;
        pp.use_e_pressure = P.use_e_pressure;
        pp.use_gravity = P.use_gravity;
        pp.initial_flow_through = P.initial_flow_through;
        pp.still = 0;
        pp.live = S_.live_src.p;
        prepare_push_fields();
        launch("push", k_push_move<Real>, tile_grid(), PM_THREADS, push_move_lds_bytes<Real>(g), g, pp,
               (const uint32_t*)pf_tile.p,
               S_.A(), S_.Dst(), (const uint32_t*)S_.sub.p, S_.appendix(), S_.cnt_tile.p,
               (const uint32_t*)S_.sub_new.p, S_.pcount.p, kbits, 0, (uint32_t)S_.cap);
        S_.pp = pp;
        S_.pushed = true;
    }

    // Merge injected particles that wait past the store's end into the
    // sorted store without pushing anything (the same two passes, still).
    void flush_injected(Species& S_) {
        if (!S_.inj_pending) return;
        if (S_.pushed) throw StateError("injected particles pending between propagate and wait_particle_sends");
        PushParams pp{};
        pp.still = 1;
        pp.live = S_.live_src.p;
        prepare_push_fields();
        launch("sort_still", k_push_move<Real>, tile_grid(), PM_THREADS, push_move_lds_bytes<Real>(g), g, pp,
               (const uint32_t*)pf_tile.p, S_.A(), S_.Dst(), (const uint32_t*)S_.sub.p, S_.appendix(), S_.cnt_tile.p,
               (const uint32_t*)S_.sub_new.p, S_.pcount.p, kbits, 0, (uint32_t)S_.cap);
        S_.pp = pp;
        S_.pushed = true;
        finish_sort(S_, "sort_still");
    }

    void scan(const uint32_t* in, uint32_t n, uint32_t* out, std::vector<DevBuf<uint32_t>>& lv, size_t level) {
        if (n <= (uint32_t)SCAN_TILE) {
            launch("scan", k_scan_apply, 1, SCAN_THREADS, 0, in, n, (const uint32_t*)zero_u32.p, out);
            return;
        }
        const uint32_t nb = cdiv(n, SCAN_TILE);
        uint32_t* bsum = lv[level].p;
        uint32_t* boff = lv[level + 1].p;
        launch("scan", k_scan_reduce, nb, SCAN_THREADS, 0, in, n, bsum);
        scan(bsum, nb, boff, lv, level + 2);
        launch("scan", k_scan_apply, nb, SCAN_THREADS, 0, in, n, (const uint32_t*)boff, out);
    }

    void wait_particle_sends(int s) override {
        SpeciesScope species_scope(timing_species, s);
        Species& S_ = S(s);
        if (!S_.pushed) throw StateError("wait_particle_sends without a preceding propagate");
        finish_sort(S_, "move");
    }

    void finish_sort(Species& S_, const char* move_name) {
        // totals and in-subcell offsets in one pass, then the scan of the totals
        launch("sort_slots", k_slot_offsets,
               N_XCD * cdiv(cdiv((uint32_t)g.ntiles * (uint32_t)g.tc, (uint32_t)SLOT_THREADS), (uint32_t)N_XCD),
               SLOT_THREADS, 0, g, S_.cnt_tile.p, (const uint8_t*)S_.live_src.p, S_.sub_new.p, dcount.p);
        scan(S_.sub_new.p, g.nkeys, S_.sub_new.p, S_.scan_lv, 0);
        launch(move_name, k_push_move<Real>, tile_grid(), PM_THREADS, move_lds_bytes(&S_), g, S_.pp,
               (const uint32_t*)pf_tile.p,
               S_.A(), S_.Dst(), (const uint32_t*)S_.sub.p, S_.appendix(), S_.cnt_tile.p,
               (const uint32_t*)S_.sub_new.p, S_.pcount.p, kbits, 1, (uint32_t)S_.cap);
        S_.swap_stores();
        S_.pushed = false;
        S_.inj_pending = false;
    }

    void clear_accumulation() override {
        for (auto& s_ : sp) s_->deposited = false;
    }

    void accumulate(int s) override {
        SpeciesScope species_scope(timing_species, s);
        Species& S_ = S(s);
        if (!S_.d.accumulate) return;
        if (S_.pushed) throw StateError("accumulate between propagate and wait_particle_sends");
        flush_injected(S_);
        const int NS = g.tc * 8;
        // tiles with at least 10 particles per subcell on average take the
        // coalesced eight-lanes-per-subcell path (measured crossover: at 8 per
        // subcell the one-lane path is faster, at 16 three times slower)
        const uint32_t dense_min = 10u * (uint32_t)NS;
        launch("deposit", k_deposit<Real>, tile_grid(), NS, deposit_lds_bytes<Real>(g), g, S_.A(),
               (const uint32_t*)S_.sub.p, (const uint16_t*)dep_list.p, (const uint8_t*)dep_count.p,
               (const uint16_t*)dep_tlist.p, (const uint8_t*)dep_tcount.p, S_.partial.p, dense_min);
        S_.deposited = true;
    }

    void add_remote_updates() override {
        const V3<Real> ji = v3(RH_CELL_JI);
        FoldSources<Real> src{};
        for (auto& s_ : sp) {
            if (!s_->deposited) continue;
            if (src.n == MAX_FOLD_SPECIES) throw ArgError("too many accumulated species");
            src.partial[src.n] = s_->partial.p;
            src.wq[src.n] = s_->d.w * s_->d.q;
            src.n++;
        }
        launch("fold", k_fold<Real>, cell_grid(), cell_block(), 0, g, src, 1.0 / (g.dx * g.dx * g.dx), P.min_rho,
               fld[RH_CELL_RHOQ].p, ji.x, ji.y, ji.z, dcount.p);
        if (ave_on)
            for (auto& s_ : sp)
                if (s_->deposited)
                    launch("average", k_average_moments<Real>, cell_grid(), cell_block(), 0, g,
                           (const Real*)s_->partial.p, s_->d.w * s_->d.q, s_->ave_n.p, s_->ave_v.p);
    }

    // ---- temporal averages (WRITE_GRID_TEMPORAL_AVERAGES) -------------------

    bool ave_on = false;
    DevBuf<double> aveB;    // running sum of cellB, SoA
    int64_t ave_count = 0;  // Hybrid::gridTemporalAverageCounter

    void set_temporal_averages(bool on) override {
        for (auto& sp_ : sp)
            if (sp_->pushed) throw StateError("temporal averages switched between propagate and wait_particle_sends");
        ave_on = on;
        ave_count = 0;
        if (!on) {
            aveB.free();
            for (auto& sp_ : sp) { sp_->ave_n.free(); sp_->ave_v.free(); }
            return;
        }
        aveB.alloc((size_t)g.nc * 3);
        fill(aveB.p, 0.0, aveB.n);
        for (auto& sp_ : sp) ensure_averages(*sp_);
        RH_HIP(hipStreamSynchronize(stream));
    }

    void ensure_averages(Species& S_) {
        if (!ave_on || S_.ave_n.p) return;
        S_.ave_n.alloc(g.nc);
        S_.ave_v.alloc((size_t)g.nc * 3);
        fill(S_.ave_n.p, 0.0, S_.ave_n.n);
        fill(S_.ave_v.p, 0.0, S_.ave_v.n);
    }

    DevBuf<float> dn_in, dn_out, dn_filt;  // denoiser volumes, filtered density (allocated on first use)
    DevBuf<uint32_t> pf_tile;     // push fields, one block per tile (k_push_fields)
    bool pf_dirty = true;         // fields changed since the blocks were written

    // The blocks of the push fields are rewritten when the fields have
    // changed; both passes of every species then read the same blocks.
    void alloc_push_fields() {
        if (pf_tile.p) return;
        pf_tile.alloc((size_t)g.ntiles * push_field_words<Real>(g));
        // halo entries of cells outside a non-periodic grid stay zero; every
        // cell's entries, its neighbourhood mask included, are written here
        // once (the fused field update then rewrites only the fields)
        launch("fill", k_fill<uint32_t>, cdiv(pf_tile.n, 256), 256, 0, pf_tile.p, 0u, pf_tile.n);
        const V3<Real> fb = v3(RH_FACE_B), ue = v3(RH_CELL_UE), ep = v3(RH_CELL_EP);
        PushFields<Real> pf{fb.x, fb.y, fb.z, ue.x, ue.y, ue.z, ep.x, ep.y, ep.z};
        launch("push_fields", k_push_fields<Real>, cell_grid(), cell_block(), 0, g, pf, (int)P.use_e_pressure,
               pf_tile.p);
    }

    void prepare_push_fields() {
        if (!pf_dirty) return;
        for (auto& s_ : sp)
            if (s_->pushed) throw StateError("fields changed between propagate and wait_particle_sends");
        alloc_push_fields();
        const V3<Real> fb = v3(RH_FACE_B), ue = v3(RH_CELL_UE), ep = v3(RH_CELL_EP);
        PushFields<Real> pf{fb.x, fb.y, fb.z, ue.x, ue.y, ue.z, ep.x, ep.y, ep.z};
        launch("push_fields", k_push_fields<Real>, cell_grid(), cell_block(), 0, g, pf, (int)P.use_e_pressure,
               pf_tile.p);
        pf_dirty = false;
    }

    void denoise(rh_denoiser* dn, double max_rel) override {
        denoise_on_stream(dn, max_rel);
        // a standalone call hands the denoiser its own stream back
        if (dn != dn_attached && rh_dn_set_stream(dn, nullptr) != RH_DN_OK)
            throw ArgError(std::string("denoiser: ") + rh_dn_last_error(dn));
    }

    void denoise_on_stream(rh_denoiser* dn, double max_rel) {
        if (!dn) throw ArgError("null denoiser");
        if (!(max_rel >= 0.0)) throw ArgError("max_rel_change must be >= 0");
        if (g.nx % 16 || g.ny % 4 || g.nz % 4) throw ArgError("denoiser needs nx % 16 == 0, ny % 4 == 0, nz % 4 == 0");
        if (!dn_in.p) {
            dn_in.alloc((size_t)g.nc * 7);
            dn_out.alloc((size_t)g.nc * 4);
            dn_filt.alloc((size_t)g.nc);
        }
        const V3<Real> ji = v3(RH_CELL_JI), b = v3(RH_CELL_B);
        launch("dn_gather", k_dn_gather<Real>, cdiv(g.nc, 256), 256, 0, g, (const Real*)fld[RH_CELL_RHOQ].p, ji, b,
               dn_in.p, dn_filt.p);
        // the denoiser's kernels are ordered on this context's stream
        if (rh_dn_set_stream(dn, (void*)stream) != RH_DN_OK ||
            rh_dn_forward(dn, dn_in.p, dn_out.p, 1, g.nx, g.ny, g.nz) != RH_DN_OK)
            throw ArgError(std::string("denoiser: ") + rh_dn_last_error(dn));
        launch("dn_scatter", k_dn_scatter<Real>, cdiv(g.nc, 256), 256, 0, g.nc, (const float*)dn_out.p,
               (const float*)dn_filt.p, max_rel, P.min_rho, fld[RH_CELL_RHOQ].p, ji, dcount.p);
    }

    rh_denoiser* dn_attached = nullptr;
    double dn_bound = 0.0;
    void set_denoiser(rh_denoiser* dn, double max_rel) override {
        if (dn && !(max_rel >= 0.0)) throw ArgError("max_rel_change must be >= 0");
        if (dn_attached && dn_attached != dn) (void)rh_dn_set_stream(dn_attached, nullptr);
        dn_attached = dn;
        dn_bound = max_rel;
    }

    void apply_bcs(int s) override {
        SpeciesScope species_scope(timing_species, s);
        Species& S_ = S(s);
        if (S_.pushed) throw StateError("boundary conditions between propagate and wait_particle_sends");
        flush_injected(S_);
        launch("boundary", k_boundary<Real>, cdiv(n_bc, BC_CELLS_PER_BLOCK), 256, 0, g, (const int*)bc_cells.p, n_bc,
               (const uint32_t*)S_.sub.p, S_.A(), S_.d.R2_obstacle, S_.bc_acc.p);
    }

    void neumann_cell(int id, int dim) {
        launch("neumann", k_neumann_cell<Real>, cdiv(n_ext, 256), 256, 0, g, (const int*)ext_cells.p, n_ext,
               fld[id].p, dim);
    }

    void cell2node(int from, int to, int dim) {
        launch("cell2node", k_cell2node<Real>, cell_grid(), cell_block(), 0, g, (const Real*)fld[from].p, fld[to].p, dim);
    }

    // The fused update (k_field_update) runs without E smoothing and when no
    // tile of it is one cell thick along a non-periodic axis (an exterior
    // cell's Neumann source then lies in its own tile).
    bool fused_field_update() const {
        if (P.n_linear_smoothings > 0 || P.use_gaussian) return false;
        const int dims[3] = {g.nx, g.ny, g.nz};
        const int ft[3] = {FieldTile<Real>::X, FieldTile<Real>::Y, FieldTile<Real>::Z};
        for (int d = 0; d < 3; ++d)
            if (!g.per[d] && dims[d] % ft[d] == 1) return false;
        return true;
    }

    void propagate_B() override { propagate_B_impl(true); }

    // full: every grid array of the reference gets its new value; otherwise
    // (the steps of rh_step before its last) only faceB, cellUe, cellEp, the
    // push-field blocks and, for an attached denoiser, cellB: the arrays the
    // next step reads.
    void propagate_B_impl(bool full) {
        for (auto& sp_ : sp)
            if (sp_->pushed) throw StateError("propagate_B between propagate and wait_particle_sends");
        if (fused_field_update()) {
            pf_dirty = true;
            alloc_push_fields();
            FieldUpdateArgs<Real> a{};
            a.faceB_in = v3(RH_FACE_B);
            a.faceB_out = {faceB_alt.p, faceB_alt.p + g.nc, faceB_alt.p + 2 * (size_t)g.nc};
            a.rho = fld[RH_CELL_RHOQ].p;
            a.Ji = v3(RH_CELL_JI);
            a.eta = fld[RH_NODE_ETA].p;
            a.inner3 = flg_fu.p;
            a.cellUe = v3(RH_CELL_UE);
            a.cellEp = v3(RH_CELL_EP);
            a.cellB = v3(RH_CELL_B);
            a.cellJ = v3(RH_CELL_J);
            a.nodeB = v3(RH_NODE_B);
            a.nodeJi = v3(RH_NODE_JI);
            a.nodeJ = v3(RH_NODE_J);
            a.nodeUe = v3(RH_NODE_UE);
            a.nodeE = v3(RH_NODE_E);
            a.nodeRho = fld[RH_NODE_RHOQ].p;
            a.push_blocks = pf_tile.p;
            a.counters = dcount.p;
            a.dt = (Real)P.dt;
            a.e_coeff = (Real)P.e_pressure_coeff;
            a.max_ue2 = (Real)P.max_electron_speed2;
            a.min_rho = (Real)P.min_rho;
            a.max_vw = (Real)P.max_whistler;
            a.max_E2 = (Real)P.max_E2;
            for (int l = 0; l < 3; ++l) { a.oue[l] = (Real)P.obstacle_ue[l]; a.imf[l] = (Real)P.imf[l]; }
            a.imf_cell = P.imf_cell[0] | P.imf_cell[1] << 1 | P.imf_cell[2] << 2 | P.imf_cell[3] << 3 |
                         P.imf_cell[4] << 4 | P.imf_cell[5] << 5;
            a.imf_face = P.imf_face[0] | P.imf_face[1] << 1 | P.imf_face[2] << 2 | P.imf_face[3] << 3 |
                         P.imf_face[4] << 4 | P.imf_face[5] << 5;
            a.use_hall = P.use_hall;
            a.use_e_pressure = P.use_e_pressure;
            a.isothermal = P.isothermal;
            a.full = full;
            a.write_cellB = dn_attached != nullptr;
            a.aveB = ave_on ? aveB.p : nullptr;
            if (ave_on) ++ave_count;
            const int nft = cdiv(g.nx, FieldTile<Real>::X) * cdiv(g.ny, FieldTile<Real>::Y) * cdiv(g.nz, FieldTile<Real>::Z);
            const dim3 grid(N_XCD * cdiv(nft, N_XCD));
            constexpr int nt = FieldUpdateThreads<Real>::N;
            if (full) launch("field_update", k_field_update<Real, true>, grid, nt, 0, g, a);
            else launch("field_update", k_field_update<Real, false>, grid, nt, 0, g, a);
            std::swap(fld[RH_FACE_B].p, faceB_alt.p);
            pf_dirty = false;
            return;
        }
        pf_dirty = true;  // faceB, cellUe and cellEp change
        const dim3 cb = cell_grid(), bb = cell_block();
        const unsigned eb = cdiv(n_ext, 256);
        const int imf_cell = P.imf_cell[0] | P.imf_cell[1] << 1 | P.imf_cell[2] << 2 | P.imf_cell[3] << 3 |
                             P.imf_cell[4] << 4 | P.imf_cell[5] << 5;
        const int imf_face = P.imf_face[0] | P.imf_face[1] << 1 | P.imf_face[2] << 2 | P.imf_face[3] << 3 |
                             P.imf_face[4] << 4 | P.imf_face[5] << 5;
        const Real bx = (Real)P.imf[0], by = (Real)P.imf[1], bz = (Real)P.imf[2];
        // face2Cell, Neumann on cellB, cellJi, cellRhoQi and the IMF on cellB
        launch("cell_B_bc", k_cell_B_bc<Real>, cb, bb, 0, g, v3(RH_FACE_B), v3(RH_CELL_B), v3(RH_CELL_JI),
               fld[RH_CELL_RHOQ].p, imf_cell, bx, by, bz);
        if (ave_on) {
            launch("average", k_average_cellB<Real>, cdiv((size_t)g.nc * 3, 256), 256, 0, g.nc,
                   (const Real*)fld[RH_CELL_B].p, aveB.p);
            ++ave_count;
        }
        // Neumann and IMF on faceB (they do not touch the cell arrays that
        // cell2Node reads), then cell2Node of cellB, cellRhoQi, cellJi and
        // calcNodeJ in one launch
        launch("neumann", k_neumann_face<Real>, eb, 256, 0, g, (const int*)ext_cells.p, n_ext, v3(RH_FACE_B));
        if (imf_face) launch("imf", k_imf<Real>, eb, 256, 0, g, (const int*)ext_cells.p, n_ext, v3(RH_FACE_B), 1, imf_face, bx, by, bz);
        {
            // the order k_node_J reads: nodeB (3), nodeRhoQi (1), nodeJi (3)
            Cell2NodeSet<Real> set{};
            const int from[3] = {RH_CELL_B, RH_CELL_RHOQ, RH_CELL_JI}, to[3] = {RH_NODE_B, RH_NODE_RHOQ, RH_NODE_JI};
            const int dim[3] = {3, 1, 3};
            for (int q = 0; q < 3; ++q) { set.from[q] = fld[from[q]].p; set.to[q] = fld[to[q]].p; set.dim[q] = dim[q]; }
            set.n = 3;
            launch("node_J", k_node_J<Real>, cb, bb, 0, g, set, v3(RH_FACE_B), v3(RH_NODE_J), (Real)P.max_whistler,
                   dcount.p);
        }
        // node2Cell of nodeJ, calcCellUe, Neumann on cellUe
        launch("cell_J_Ue", k_cell_J_Ue<Real>, cb, bb, 0, g, v3(RH_NODE_J), v3(RH_CELL_J), v3(RH_CELL_JI),
               fld[RH_CELL_RHOQ].p, (const uint8_t*)flg[RH_INNER_FIELD].p, v3(RH_CELL_UE), P.use_hall,
               (Real)P.obstacle_ue[0], (Real)P.obstacle_ue[1], (Real)P.obstacle_ue[2], (Real)P.max_electron_speed2,
               (Real)P.min_rho, dcount.p);
        // calcNodeUe, upwindNodeB, calcNodeE
        launch("node_Ue_B_E", k_node_Ue_B_E<Real>, cb, bb, 0, g, v3(RH_NODE_J), v3(RH_NODE_JI), fld[RH_NODE_RHOQ].p,
               (const uint8_t*)flg[RH_INNER_NODE].p, v3(RH_NODE_UE), P.use_hall, (Real)P.obstacle_ue[0],
               (Real)P.obstacle_ue[1], (Real)P.obstacle_ue[2], (Real)P.max_electron_speed2, (Real)P.min_rho,
               v3(RH_CELL_B), v3(RH_NODE_B), (const Real*)fld[RH_NODE_ETA].p, v3(RH_NODE_E), (Real)P.max_E2,
               dcount.p);
        for (int it = 0; it < P.n_linear_smoothings; ++it) {
            launch("node2cell", k_node2cell<Real>, cb, bb, 0, g, v3(RH_NODE_E), v3(RH_CELL_J));
            neumann_cell(RH_CELL_J, 3);
            cell2node(RH_CELL_J, RH_NODE_E, 3);
            fill(fld[RH_CELL_J].p, Real(0), fld[RH_CELL_J].n);
        }
        if (P.use_gaussian) {
            RH_HIP(hipMemcpyAsync(tmpv.p, fld[RH_NODE_E].p, sizeof(Real) * 3 * g.nc, hipMemcpyDeviceToDevice, stream));
            V3<Real> old{tmpv.p, tmpv.p + g.nc, tmpv.p + 2 * (size_t)g.nc};
            launch("node_avg", k_node_avg<Real>, cb, bb, 0, g, old, v3(RH_NODE_E), (Real)P.gauss_coeffs[0],
                   (Real)P.gauss_coeffs[1], (Real)P.gauss_coeffs[2], (Real)P.gauss_coeffs[3]);
        }
        // faceCurl, calcCellEp and the next push's per-tile field blocks
        alloc_push_fields();
        launch("face_curl_Ep", k_face_curl_Ep<Real>, cb, bb, 0, g, v3(RH_NODE_E), v3(RH_FACE_B), (Real)P.dt,
               (int)P.use_e_pressure, (const Real*)fld[RH_NODE_RHOQ].p, (const Real*)fld[RH_CELL_RHOQ].p,
               (const uint8_t*)flg[RH_INNER_CELLEP].p, v3(RH_CELL_EP), (Real)P.e_pressure_coeff, P.isothermal,
               v3(RH_CELL_UE), pf_tile.p);
        pf_dirty = false;
    }

    void step(int n) override {
        for (int it = 0; it < n; ++it) {
            ++tstep;
            if (species_streams && sp.size() > 1) {
                particle_phase_side_by_side();
            } else {
                for (size_t s = 0; s < sp.size(); ++s) propagate((int)s);
                clear_accumulation();
                for (size_t s = 0; s < sp.size(); ++s) wait_particle_sends((int)s);
                for (size_t s = 0; s < sp.size(); ++s) accumulate((int)s);
            }
            add_remote_updates();
            for (size_t s = 0; s < sp.size(); ++s) apply_bcs((int)s);
            for (size_t s = 0; s < sp.size(); ++s) inject((int)s);
            if (dn_attached) denoise_on_stream(dn_attached, dn_bound);
            // only the last step's update writes the arrays nothing in
            // between reads
            propagate_B_impl(it == n - 1);
        }
        check_overflow();
        // each species' particle count (the sorted store's end; the stream has
        // finished, check_overflow waited for it), for the next call's choice
        // of the move's occupancy
        for (auto& s_ : sp) {
            uint32_t np = 0;
            RH_HIP(hipMemcpyAsync(&np, s_->sub.p + g.nkeys, sizeof(np), hipMemcpyDeviceToHost, stream));
            RH_HIP(hipStreamSynchronize(stream));
            s_->dense = (double)np >= 4.0 * (double)g.nc;
        }
    }

    // propagate, wait_particle_sends and accumulate of every species, each
    // species on a stream of its own after the work before them on the
    // context's stream, which then waits for all of them.  The species share
    // only read-only inputs (the fields and the push blocks, written before
    // the fork) and integer counters; every buffer they write is their own,
    // so the results are bit for bit those of the one-stream order.  A
    // species whose few crowded tiles keep a handful of CUs busy while the
    // rest of the device idles (photoions piled up near an obstacle) then
    // runs beside the other species' full-device launches.
    struct LaunchStream {
        hipStream_t& slot;
        hipStream_t prev;
        LaunchStream(hipStream_t& s, hipStream_t to) : slot(s), prev(s) { slot = to; }
        ~LaunchStream() { slot = prev; }
    };
    void particle_phase_side_by_side() {
        for (auto& s_ : sp)
            if (s_->pushed) throw StateError("step called between propagate and wait_particle_sends");
        prepare_push_fields();
        clear_accumulation();
        RH_HIP(hipEventRecord(fork_ev, stream));
        for (size_t s = 0; s < sp.size(); ++s) {
            Species& S_ = *sp[s];
            if (!S_.ss) {
                RH_HIP(hipStreamCreateWithFlags(&S_.ss, hipStreamNonBlocking));
                RH_HIP(hipEventCreateWithFlags(&S_.done, hipEventDisableTiming));
            }
            RH_HIP(hipStreamWaitEvent(S_.ss, fork_ev, 0));
            {
                LaunchStream on(ls, S_.ss);
                propagate((int)s);
                wait_particle_sends((int)s);
                accumulate((int)s);
            }
            RH_HIP(hipEventRecord(S_.done, S_.ss));
        }
        for (auto& s_ : sp) RH_HIP(hipStreamWaitEvent(stream, s_->done, 0));
    }

    // The sort and the injection count overflows of a particle store into
    // the overflow counter; its copies in every counter slot are summed, so
    // that the steps that lost particles report it whichever slot a kernel
    // used.
    void check_overflow() {
        unsigned long long h[C_N];
        counter_totals(h);
        if (h[C_RANK_OVERFLOW])
            throw CapError("particle sort overflow (a subcell holds more than 65535 particles of one species, "
                           "or the particle store is full)");
    }

    // ---- injection -------------------------------------------------------

    int64_t tstep = 0;  // Simulation::timestep: advanced by one at the start of every step
    void set_timestep(int64_t t) override { tstep = t; }
    int64_t timestep() override { return tstep; }

    void set_injector(int s, const rh_injector_desc* d) override {
        Species& S_ = S(s);
        if (S_.pushed) throw StateError("set_injector between propagate and wait_particle_sends");
        flush_injected(S_);
        carry_acc(S_, false, true);
        S_.inj = InjParams{};
        S_.inj_ncell = 0;
        S_.inj_slots = 0;
        if (!d || d->kind == RH_INJ_NONE) return;
        InjParams ip{};
        ip.k0 = (uint32_t)d->seed;
        ip.k1 = (uint32_t)(d->seed >> 32);
        if (!(d->temperature > 0)) throw ArgError("injector temperature must be positive");
        ip.vth = std::sqrt(KB * d->temperature / S_.d.m);
        std::vector<int> cells;
        std::vector<double> rate;
        if (d->kind == RH_INJ_SOLAR_WIND) {
            // InjectorSolarWind::initialize (src/particle_injector.cpp:588-682)
            int nzc = 0, axis = -1;
            for (int l = 0; l < 3; ++l)
                if (d->velocity[l] != 0.0) { ++nzc; axis = l; }
            if (nzc != 1) throw ArgError("solar wind velocity must have exactly one non-zero component");
            if (!(d->ppc > 0) || !(d->density > 0)) throw ArgError("solar wind density and ppc must be positive");
            const double w = d->density * g.dx * g.dx * g.dx / d->ppc;
            if (!(std::fabs(w - S_.d.w) <= 1e-9 * std::fabs(w)))
                throw ArgError("species weight differs from density*dV/ppc of its solar wind injector");
            const double U = std::fabs(d->velocity[axis]);
            ip.kind = INJ_SOLAR_WIND;
            ip.axis = axis;
            ip.pos_wall = d->velocity[axis] < 0;
            ip.rate = d->ppc * U * P.dt / g.dx;
            ip.x0 = U / ip.vth;
            deriv_gauss_consts(ip.x0, ip.xm, ip.c, ip.d, ip.cxm);
            // checkIfInjectionCell{X,Y,Z}{Pos,Neg} (:392-454): the wall
            // neighbour is missing, the four transverse neighbours exist
            const int dims[3] = {g.nx, g.ny, g.nz};
            for (int n = 0; n < g.nc; ++n) {
                int c3[3];
                cell_ijk(g, n, c3[0], c3[1], c3[2]);
                bool ok = ip.pos_wall ? !has_pos(c3[axis], dims[axis], g.per[axis]) : !has_neg(c3[axis], g.per[axis]);
                for (int l = 0; l < 3 && ok; ++l)
                    if (l != axis) ok = has_pos(c3[l], dims[l], g.per[l]) && has_neg(c3[l], g.per[l]);
                if (ok) cells.push_back(n);
            }
        } else if (d->kind == RH_INJ_UNIFORM) {
            // InjectorUniform::initialize (src/particle_injector.cpp:267-345)
            if (!(d->ppc > 0) || !(d->density > 0)) throw ArgError("uniform density and ppc must be positive");
            const double w = d->density * g.dx * g.dx * g.dx / d->ppc;
            if (!(std::fabs(w - S_.d.w) <= 1e-9 * std::fabs(w)))
                throw ArgError("species weight differs from density*dV/ppc of its uniform injector");
            for (int l = 0; l < 3; ++l)
                if (!(d->box[2 * l] < d->box[2 * l + 1])) throw ArgError("uniform injector box: min >= max");
            ip.kind = INJ_UNIFORM;
            ip.rate = d->ppc;
            ip.eps = 1.0e-4;  // epsCell (src/particle_injector.cpp:31)
            for (int l = 0; l < 3; ++l) ip.U[l] = d->velocity[l];
            const double o3[3] = {g.x0, g.y0, g.z0};
            for (int n = 0; n < g.nc; ++n) {
                int c3[3];
                cell_ijk(g, n, c3[0], c3[1], c3[2]);
                bool in = true;
                for (int l = 0; l < 3; ++l) {
                    const double c = o3[l] + (c3[l] + 0.5) * g.dx;
                    in = in && !(c < d->box[2 * l] || c > d->box[2 * l + 1]);
                }
                if (in) cells.push_back(n);
            }
        } else if (d->kind == RH_INJ_EXOSPHERE) {
            if (!d->cell_rate) throw ArgError("exosphere injector needs cell_rate");
            ip.kind = INJ_EXOSPHERE;
            ip.eps = 1.0e-4;  // epsCell (src/particle_injector.cpp:31)
            ip.R2_exobase = d->exobase_radius > 0 ? d->exobase_radius * d->exobase_radius : 0.0;
            for (int n = 0; n < g.nc; ++n) {
                const double r = d->cell_rate[n];
                if (!(r >= 0) || !std::isfinite(r)) throw ArgError("cell_rate must be finite and non-negative");
                if (r > 0) cells.push_back(n);
            }
        } else {
            throw ArgError("unknown injector kind");
        }
        // listed cells in subcell-key order: each tile's cells are one range
        std::vector<uint32_t> key(cells.size());
        for (size_t q = 0; q < cells.size(); ++q) {
            int i, j, k;
            cell_ijk(g, cells[q], i, j, k);
            key[q] = cell_key(g, i, j, k);
        }
        std::vector<size_t> order(cells.size());
        for (size_t q = 0; q < order.size(); ++q) order[q] = q;
        std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return key[a] < key[b]; });
        std::vector<int> hc(cells.size());
        std::vector<double> hr(cells.size());
        std::vector<uint32_t> cbeg((size_t)g.ntiles + 1, 0);
        for (size_t q = 0; q < order.size(); ++q) {
            hc[q] = cells[order[q]];
            if (ip.kind == INJ_EXOSPHERE) hr[q] = d->cell_rate[cells[order[q]]];
            cbeg[key[order[q]] / (uint32_t)g.tc + 1]++;
        }
        for (int T = 0; T < g.ntiles; ++T) cbeg[T + 1] += cbeg[T];
        const size_t nl = std::max<size_t>(hc.size(), 1);
        S_.inj_cells.alloc(nl);
        S_.inj_rate.alloc(ip.kind == INJ_EXOSPHERE ? nl : 1);
        S_.inj_cbeg.alloc(cbeg.size());
        S_.inj_off.alloc(nl + 1);
        if (!S_.inj_total.p) {
            S_.inj_total.alloc(1);
            RH_HIP(hipMemsetAsync(S_.inj_total.p, 0, sizeof(unsigned long long), stream));
        }
        if (!hc.empty()) {
            RH_HIP(hipMemcpyAsync(S_.inj_cells.p, hc.data(), hc.size() * sizeof(int), hipMemcpyHostToDevice, stream));
            if (ip.kind == INJ_EXOSPHERE)
                RH_HIP(hipMemcpyAsync(S_.inj_rate.p, hr.data(), hr.size() * sizeof(double), hipMemcpyHostToDevice, stream));
        }
        RH_HIP(hipMemcpyAsync(S_.inj_cbeg.p, cbeg.data(), cbeg.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        int slots = 0;
        std::vector<uint32_t> off;
        if (ip.kind == INJ_SOLAR_WIND || ip.kind == INJ_UNIFORM) {
            // fixed places: probRound never exceeds ceil(rate)
            const double cr = std::ceil(ip.rate);
            if (cr * (double)std::max<size_t>(hc.size(), 1) >= 2147483647.0)
                throw ArgError("injection rate too large for the fixed places");
            slots = (int)cr;
            off.resize(hc.size() + 1);
            for (size_t q = 0; q <= hc.size(); ++q) off[q] = (uint32_t)(q * (size_t)slots);
            RH_HIP(hipMemcpyAsync(S_.inj_off.p, off.data(), off.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        }
        // inject energy accumulators: per block of places (fixed places), per listed cell (count and fill)
        S_.inj_acc.alloc(slots > 0 ? std::max<size_t>(cdiv(hc.size() * (size_t)slots, 256), 1) : nl);
        fill(S_.inj_acc.p, 0.0, S_.inj_acc.n);
        RH_HIP(hipStreamSynchronize(stream));
        S_.inj = ip;
        S_.inj_ncell = (int)hc.size();
        S_.inj_slots = slots;
    }

    // ParticleInjectorBase::inject for the current timestep (nothing at
    // timestep <= 0, src/particle_injector.cpp:544): count, scan, fill.
    void inject(int s) override {
        SpeciesScope species_scope(timing_species, s);
        Species& S_ = S(s);
        if (S_.pushed) throw StateError("inject between propagate and wait_particle_sends");
        if (S_.inj.kind == INJ_NONE || S_.inj_ncell == 0 || tstep <= 0) return;
        if (S_.inj.kind == INJ_UNIFORM && tstep != 1) return;  // src/particle_injector.cpp:181
        flush_injected(S_);  // a second injection before the next push
        InjParams ip = S_.inj;
        ip.step = (uint32_t)tstep;
        if (ip.kind == INJ_SOLAR_WIND || ip.kind == INJ_UNIFORM) {
            if (S_.inj_slots > 0)
                launch("inject_slots", k_inject_slots<Real>, cdiv((size_t)S_.inj_ncell * S_.inj_slots, 256), 256, 0, g,
                       ip, S_.inj_ncell, S_.inj_slots, (const int*)S_.inj_cells.p,
                       (const uint32_t*)(S_.sub.p + g.nkeys), S_.A(), (uint32_t)S_.cap, S_.inj_total.p, dcount.p,
                       S_.inj_acc.p);
            S_.inj_pending = S_.inj_slots > 0;
            return;
        }
        const unsigned nb = cdiv(S_.inj_ncell, 256);
        const double* rate = S_.inj_rate.p;
        launch("inject_count", k_inject_count<Real>, nb, 256, 0, g, ip, S_.inj_ncell, (const int*)S_.inj_cells.p,
               rate, S_.inj_off.p);
        scan(S_.inj_off.p, (uint32_t)S_.inj_ncell, S_.inj_off.p, scan_lv, 0);
        launch("inject_fill", k_inject_fill<Real>, nb, 256, 0, g, ip, S_.inj_ncell, (const int*)S_.inj_cells.p, rate,
               (const uint32_t*)S_.inj_off.p, (const uint32_t*)(S_.sub.p + g.nkeys), S_.A(), (uint32_t)S_.cap,
               S_.inj_total.p, dcount.p, S_.inj_acc.p);
        S_.inj_pending = true;
    }

    // ---- logs (logWriteParticleField, src/diagnostics.h:407-813) -------------

    double log_t0 = 0.0;                 // Hybrid::logCounterTimeStart
    unsigned long long log_cnt0[C_N] = {};
    DevBuf<double> log_rhoM, log_partial;
    std::ofstream log_field;
    std::vector<std::unique_ptr<std::ofstream>> log_pop;
    int log_precision = 10, log_inner = 0;
    double log_max_B = 0.0;

    // FP64 column sums of a rows x cols accumulator, in a fixed order.
    void sum_cols(const double* a, size_t rows, int cols, double* out) {
        if (!log_partial.p || log_partial.n < (size_t)LOG_BLOCKS * LQ_N_Q) log_partial.alloc((size_t)LOG_BLOCKS * LQ_N_Q);
        for (int c = 0; c < cols; ++c) out[c] = 0.0;
        if (rows == 0) return;
        launch("log_sum", k_sum_cols, LOG_BLOCKS, 256, 0, a, rows, cols, log_partial.p);
        std::vector<double> h((size_t)LOG_BLOCKS * cols);
        RH_HIP(hipMemcpyAsync(h.data(), log_partial.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, stream));
        RH_HIP(hipStreamSynchronize(stream));
        for (int b = 0; b < LOG_BLOCKS; ++b)
            for (int c = 0; c < cols; ++c) out[c] += h[(size_t)b * cols + c];
    }

    // Move the boundary / injection accumulators into the host carries
    // (before their layout changes) and clear them.
    void carry_acc(Species& S_, bool bc, bool inj) {
        if (bc && S_.bc_acc.p) {
            double t[4];
            sum_cols(S_.bc_acc.p, (size_t)n_bc, 4, t);
            for (int c = 0; c < 4; ++c) S_.carry_bc[c] += t[c];
            fill(S_.bc_acc.p, 0.0, S_.bc_acc.n);
        }
        if (inj && S_.inj_acc.p) {
            double t = 0.0;
            sum_cols(S_.inj_acc.p, S_.inj_acc.n, 1, &t);
            S_.carry_inj_v2 += t;
            fill(S_.inj_acc.p, 0.0, S_.inj_acc.n);
        }
    }

    void log_begin(double t) override {
        log_t0 = t;
        counter_totals(log_cnt0);
        for (auto& sp_ : sp) {
            Species& S_ = *sp_;
            if (S_.bc_acc.p) fill(S_.bc_acc.p, 0.0, S_.bc_acc.n);
            if (S_.inj_acc.p) fill(S_.inj_acc.p, 0.0, S_.inj_acc.n);
            for (int c = 0; c < 4; ++c) S_.carry_bc[c] = 0.0;
            S_.carry_inj_v2 = 0.0;
            S_.log_maxvi0 = (double)species_counter(S_, C_MAX_VI);
            S_.log_inj0 = injected_count(S_);
        }
        RH_HIP(hipStreamSynchronize(stream));
    }

    int64_t injected_count(Species& S_) {
        if (!S_.inj_total.p) return 0;
        unsigned long long h = 0;
        RH_HIP(hipMemcpyAsync(&h, S_.inj_total.p, sizeof(h), hipMemcpyDeviceToHost, stream));
        RH_HIP(hipStreamSynchronize(stream));
        return (int64_t)h;
    }

    // One line of every pop log (17 values per species) and of field.log
    // (41 values, src/user.cpp:2161-2231 list the columns); the counters
    // restart at t.  Column 31 repeats the node-Ue counter, as
    // src/diagnostics.h:589 does.
    void log_values(double t, int include_inner, double* field, double* pop) override {
        for (auto& sp_ : sp)
            if (sp_->pushed) throw StateError("log between propagate and wait_particle_sends");
        for (auto& sp_ : sp) flush_injected(*sp_);
        const double dt = P.dt, dV = g.dx * g.dx * g.dx, Dt = t - log_t0;
        if (!log_rhoM.p) log_rhoM.alloc((size_t)g.nc);
        if (sp.empty()) fill(log_rhoM.p, 0.0, log_rhoM.n);
        unsigned long long cnt[C_N];
        counter_totals(cnt);
        for (size_t s = 0; s < sp.size(); ++s) {
            Species& S_ = *sp[s];
            double d[DIAG_Q];
            diag_raw((int)s, d);
            launch("log_mass", k_cell_mass<Real>, cdiv(g.nc, 256), 256, 0, g, (const uint32_t*)S_.sub.p, S_.A(),
                   S_.d.m * S_.d.w, log_rhoM.p, s == 0 ? 1 : 0);
            double bc[4], v2inj = 0.0;
            sum_cols(S_.bc_acc.p, (size_t)n_bc, 4, bc);
            for (int c = 0; c < 4; ++c) bc[c] += S_.carry_bc[c];
            if (S_.inj_acc.p) sum_cols(S_.inj_acc.p, S_.inj_acc.n, 1, &v2inj);
            v2inj += S_.carry_inj_v2;
            const double ninj = (double)(injected_count(S_) - S_.log_inj0);
            const double nvi = (double)species_counter(S_, C_MAX_VI) - S_.log_maxvi0;
            const double w = S_.d.w, m = S_.d.m, Nr = d[0] * w;
            double* o = pop + 17 * s;
            o[0] = t; o[1] = Nr; o[2] = d[0];
            if (Nr > 0) { o[3] = d[1] * w / Nr; o[4] = d[2] * w / Nr; o[5] = d[3] * w / Nr; o[6] = d[4] * w / Nr; }
            else { o[3] = o[4] = o[5] = o[6] = 0.0; }
            o[7] = 0.5 * m * (d[5] * w); o[8] = d[6];
            if (Dt > 0) {
                o[9] = bc[0] * w / Dt; o[10] = bc[2] * w / Dt; o[11] = ninj * w / Dt; o[12] = ninj / Dt * dt;
                o[13] = 0.5 * m * (bc[1] * w) / Dt; o[14] = 0.5 * m * (bc[3] * w) / Dt;
                o[15] = 0.5 * m * (v2inj * w) / Dt; o[16] = nvi / Dt * dt;
            } else {
                for (int q = 9; q < 17; ++q) o[q] = 0.0;
            }
        }
        if (!log_partial.p || log_partial.n < (size_t)LOG_BLOCKS * LQ_N_Q) log_partial.alloc((size_t)LOG_BLOCKS * LQ_N_Q);
        LogFields<Real> lf{v3(RH_FACE_B), v3(RH_CELL_JI), v3(RH_CELL_EP), v3(RH_NODE_E), v3(RH_CELL_UE),
                           fld[RH_CELL_RHOQ].p, flg[RH_INNER_FIELD].p, log_rhoM.p};
        const double min_rhoM = P.min_rho * LOG_MP / LOG_E;
        launch("log_field", k_field_log<Real>, LOG_BLOCKS, 256, 0, g, lf, include_inner, min_rhoM, log_partial.p);
        std::vector<double> h((size_t)LOG_BLOCKS * LQ_N_Q);
        RH_HIP(hipMemcpyAsync(h.data(), log_partial.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, stream));
        RH_HIP(hipStreamSynchronize(stream));
        double v[LQ_N_Q];
        for (int q = 0; q < LQ_N_Q; ++q) v[q] = h[q];
        for (int b = 1; b < LOG_BLOCKS; ++b)
            for (int q = 0; q < LQ_N_Q; ++q) {
                const double x = h[(size_t)b * LQ_N_Q + q];
                v[q] = q < LQ_NSUM ? v[q] + x : (q < LQ_NMAX_END ? std::max(v[q], x) : std::min(v[q], x));
            }
        const double N = v[LQ_N];
        auto avg = [&](int q) { return N > 0 ? v[q] / N : 0.0; };
        field[0] = t;
        field[1] = avg(LQ_BX); field[2] = avg(LQ_BY); field[3] = avg(LQ_BZ); field[4] = avg(LQ_BT);
        field[5] = v[LQ_MAX_B]; field[6] = avg(LQ_DIV); field[7] = v[LQ_MAX_DIV];
        field[8] = g.dx * v[LQ_MAX_DIVPB]; field[9] = v[LQ_B2] * dV / (2.0 * LOG_MU0);
        field[10] = avg(LQ_JX); field[11] = avg(LQ_JY); field[12] = avg(LQ_JZ); field[13] = avg(LQ_JT);
        field[14] = v[LQ_MAX_J];
        field[15] = avg(LQ_EPX); field[16] = avg(LQ_EPY); field[17] = avg(LQ_EPZ); field[18] = avg(LQ_EPT);
        field[19] = v[LQ_MAX_EP]; field[20] = v[LQ_EP2] * dV * LOG_EPS0 / 2.0;
        field[21] = avg(LQ_EX); field[22] = avg(LQ_EY); field[23] = avg(LQ_EZ); field[24] = avg(LQ_ET);
        field[25] = v[LQ_MAX_E]; field[26] = v[LQ_E2] * dV * LOG_EPS0 / 2.0; field[27] = v[LQ_CE2] * dV * LOG_EPS0 / 2.0;
        const int ids[6] = {C_MAX_CELL_UE, C_MAX_NODE_UE, C_MAX_NODE_UE, C_MAX_E, C_MIN_CELL_RHO, C_MIN_NODE_RHO};
        for (int q = 0; q < 6; ++q) field[28 + q] = Dt > 0 ? (double)(cnt[ids[q]] - log_cnt0[ids[q]]) / Dt * dt : 0.0;
        field[34] = v[LQ_MIN_DE]; field[35] = v[LQ_MAX_DE]; field[36] = v[LQ_MIN_DI]; field[37] = v[LQ_MAX_DI];
        field[38] = v[LQ_MIN_TL]; field[39] = v[LQ_MAX_VA]; field[40] = v[LQ_MAX_UE];
        log_begin(t);
    }

    // The files and headers of src/user.cpp:2147-2231.
    void log_open(const char* dir, const char* const* names, int precision, int include_inner,
                  double max_B_stop) override {
        if (!dir || !names) throw ArgError("log directory and species names are required");
        if (precision < 1) precision = 1;
        log_precision = precision;
        log_inner = include_inner;
        log_max_B = max_B_stop;
        const std::string d(dir);
        log_pop.clear();
        for (size_t s = 0; s < sp.size(); ++s) {
            if (!names[s]) throw ArgError("null species name");
            // the zero padding follows the 0-based index and the number is
            // 1-based, as in the reference
            const char* zero = s < 10 ? "00" : (s < 100 ? "0" : "");
            std::ostringstream fn;
            fn << d << "/pop" << zero << s + 1 << "_" << names[s] << ".log";
            auto f = std::make_unique<std::ofstream>(fn.str().c_str(), std::ios_base::out);
            if (!*f) throw ArgError("cannot open " + fn.str());
            f->precision(precision);
            *f << std::scientific << std::showpos;
            *f << "% " << names[s] << std::endl
               << "% m [kg] = " << sp[s]->d.m << std::endl
               << "% q [C] = " << sp[s]->d.q << std::endl
               << "% columns = 17" << std::endl
               << "% 01. Time [s]" << std::endl << "% 02. Particles [#]" << std::endl
               << "% 03. Macroparticles [#]" << std::endl << "% 04. avg(Vx) [m/s]" << std::endl
               << "% 05. avg(Vy) [m/s]" << std::endl << "% 06. avg(Vz) [m/s]" << std::endl
               << "% 07. avg(|V|) [m/s]" << std::endl << "% 08. Kinetic energy [J]" << std::endl
               << "% 09. max(|V|) [m/s]" << std::endl << "% 10. Escape rate [#/s]" << std::endl
               << "% 11. Impact rate [#/s]" << std::endl << "% 12. Inject rate [#/s]" << std::endl
               << "% 13. Macroparticle inject rate [#/dt]" << std::endl
               << "% 14. Kinetic energy escape rate [J/s]" << std::endl
               << "% 15. Kinetic energy impact rate [J/s]" << std::endl
               << "% 16. Kinetic energy inject rate [J/s]" << std::endl
               << "% 17. Constraint maxVi rate [#/dt]" << std::endl;
            log_pop.push_back(std::move(f));
        }
        if (log_field.is_open()) log_field.close();
        log_field.open((d + "/field.log").c_str(), std::ios_base::out);
        if (!log_field) throw ArgError("cannot open " + d + "/field.log");
        log_field.precision(precision);
        log_field << std::scientific << std::showpos;
        static const char* cols[41] = {
            "Time [s]", "avg(faceBx) [T]", "avg(faceBy) [T]", "avg(faceBz) [T]", "avg(|faceB|) [T]", "max(|faceB|) [T]",
            "avg(div(faceB)) [T/m]", "max(div(faceB)) [T/m]", "max(dx*div(faceB)/faceB) [-]",
            "energy(sum(dV*faceB^2/2*mu0)) [J]", "avg(cellJix) [A/m^2]", "avg(cellJiy) [A/m^2]",
            "avg(cellJiz) [A/m^2]", "avg(|cellJi|) [A/m^2]", "max(|cellJi|) [A/m^2]", "avg(cellEpx) [V/m]",
            "avg(cellEpy) [V/m]", "avg(cellEpz) [V/m]", "avg(|cellEp|) [V/m]", "max(|cellEp|) [V/m]",
            "energy(sum(dV*cellEp^2*eps0/2)) [J]", "avg(nodeEx) [V/m]", "avg(nodeEy) [V/m]", "avg(nodeEz) [V/m]",
            "avg(|nodeE|) [V/m]", "max(|nodeE|) [V/m]",
            "energy(sum(dV*nodeE^2*eps0/2)), nodeE = -Ue x B + eta*J [J]",
            "energy(sum(dV*cellE^2*eps0/2)), cellE = -Ue x B [J]", "Constraint maxUe rate (cell) [#/dt]",
            "Constraint maxUe rate (node) [#/dt]", "Constraint maxVw rate (node) [#/dt]",
            "Constraint maxE rate (node) [#/dt]", "Constraint minRhoQi rate (cell) [#/dt]",
            "Constraint minRhoQi rate (node) [#/dt]", "Minimum electron inertial length [m]",
            "Maximum electron inertial length [m]", "Minimum proton inertial length [m]",
            "Maximum proton inertial length [m]", "Minimum ion Larmor period [s]", "Maximum Alfven speed [m/s]",
            "Maximum electron fluid speed [m/s]"};
        log_field << "% field" << std::endl << "% columns = 41" << std::endl;
        for (int c = 0; c < 41; ++c) {
            char num[8];
            std::snprintf(num, sizeof(num), "%02d", c + 1);  // the stream shows signs
            log_field << "% " << num << ". " << cols[c] << std::endl;
        }
    }

    void log_write(double t) override {
        if (!log_field.is_open()) throw StateError("log files are not open (rh_log_open)");
        std::vector<double> field(41), pop(17 * sp.size());
        log_values(t, log_inner, field.data(), pop.data());
        for (size_t s = 0; s < sp.size(); ++s) {
            std::ofstream& f = *log_pop[s];
            for (int q = 0; q < 17; ++q) f << pop[17 * s + q] << " ";
            f << std::endl;
        }
        // the reference writes max(|nodeE|) after an extra space
        for (int q = 0; q < 41; ++q) log_field << (q == 25 ? " " : "") << field[q] << " ";
        log_field << std::endl;
        if (log_max_B > 0 && field[5] > log_max_B)
            throw StateError("maximum B reached: the reference stops the run here (max_magnetic_field_stop_run)");
    }

    // ---- VLSV output (UserDataOP::writeData, src/operator_userdata.cpp:55-311)

    // The requested cell variables, in the reference's order, on the mesh
    // "SpatialGrid" (cells n = i + nx*(j + ny*k), CellID = n), plus the mesh
    // description the reference's reader uses (MESH_BBOX, node coordinates,
    // CellID, parameters).  pops[s] names the output population of species s
    // (species with one name are combined, as Hybrid::outputPopVarStr groups
    // them) for the n_, v_ and T_ variables.
    void write_vlsv(const char* path, double t, int64_t timestep, const char* const* vars, int nvars,
                    const char* const* pops) override {
        if (!path || (nvars > 0 && !vars)) throw ArgError("null VLSV path or variable list");
        for (auto& sp_ : sp)
            if (sp_->pushed) throw StateError("VLSV output between propagate and wait_particle_sends");
        std::map<std::string, bool> want;
        for (int q = 0; q < nvars; ++q) {
            if (!vars[q]) throw ArgError("null VLSV variable name");
            want[vars[q]] = true;
        }
        static const char* known[] = {"faceB", "cellRhoQi", "cellB", "cellJ", "cellUe", "cellJi", "cellEp",
                                      "nodeRhoQi", "nodeE", "nodeB", "nodeJ", "nodeUe", "nodeJi", "nodeEta",
                                      "innerFlagField", "innerFlagNode", "innerFlagParticle", "innerFlagCellEp",
                                      "ne", "cellDivB", "cellNPles", "n", "v", "T", "cellBAverage", "n_ave",
                                      "v_ave"};
        for (const auto& kv : want) {
            bool ok = false;
            for (const char* k : known) ok = ok || kv.first == k;
            if (!ok) throw ArgError("unknown VLSV variable " + kv.first);
        }
        if ((want.count("n") || want.count("v") || want.count("T") || want.count("n_ave") || want.count("v_ave")) &&
            !pops)
            throw ArgError("population names are required for n, v, T, n_ave and v_ave");
        if ((want.count("cellBAverage") || want.count("n_ave") || want.count("v_ave")) && !ave_on)
            throw ArgError("temporal averages are off (rh_set_temporal_averages)");
        for (auto& sp_ : sp) flush_injected(*sp_);
        const uint64_t nc = (uint64_t)g.nc;
        VlsvWriter W(path);
        vlsv_write_grid(W, VlsvGrid{g.nx, g.ny, g.nz, g.dx, g.x0, g.y0, g.z0}, t, timestep, P.dt);
        auto var = [&](const std::string& name, const void* data, uint64_t vec, const char* type, uint64_t size) {
            vlsv_write_cell_var(W, nc, name, data, vec, type, size);
        };
        std::vector<double> h;
        auto field_var = [&](const char* name, int id) {
            if (!want.count(name)) return;
            h.assign(nc * FIELD_DIM[id], 0.0);
            download_field(id, h.data());
            var(name, h.data(), FIELD_DIM[id], "float", 8);
        };
        field_var("faceB", RH_FACE_B);
        field_var("cellRhoQi", RH_CELL_RHOQ);
        field_var("cellB", RH_CELL_B);
        field_var("cellJ", RH_CELL_J);
        field_var("cellUe", RH_CELL_UE);
        field_var("cellJi", RH_CELL_JI);
        if (P.use_e_pressure) field_var("cellEp", RH_CELL_EP);
        field_var("nodeRhoQi", RH_NODE_RHOQ);
        field_var("nodeE", RH_NODE_E);
        field_var("nodeB", RH_NODE_B);
        field_var("nodeJ", RH_NODE_J);
        field_var("nodeUe", RH_NODE_UE);
        field_var("nodeJi", RH_NODE_JI);
        field_var("nodeEta", RH_NODE_ETA);
        const char* flag_names[RH_N_FLAGS] = {"innerFlagField", "innerFlagNode", "innerFlagParticle", "innerFlagCellEp"};
        for (int f = 0; f < RH_N_FLAGS; ++f) {
            if (!want.count(flag_names[f])) continue;
            std::vector<uint8_t> b(nc);
            RH_HIP(hipMemcpyAsync(b.data(), flg[f].p, nc, hipMemcpyDeviceToHost, stream));
            RH_HIP(hipStreamSynchronize(stream));
            var(flag_names[f], b.data(), 1, "uint", 1);  // bool arrays, one byte each
        }
        if (want.count("ne")) {
            h.assign(nc, 0.0);
            download_field(RH_CELL_RHOQ, h.data());
            for (auto& x : h) x /= LOG_E;
            var("ne", h.data(), 1, "float", 8);
        }
        if (want.count("cellDivB")) {
            // calcCellDiv (:338-352): cells with all 26 neighbours, 0 elsewhere
            std::vector<double> fb(nc * 3), dv(nc, 0.0);
            download_field(RH_FACE_B, fb.data());
            for (int n = 0; n < g.nc; ++n) {
                int i, j, k;
                cell_ijk(g, n, i, j, k);
                if (!all_nbrs(g, i, j, k)) continue;
                const int xm = nbr(g, i, j, k, -1, 0, 0), ym = nbr(g, i, j, k, 0, -1, 0), zm = nbr(g, i, j, k, 0, 0, -1);
                dv[n] = ((fb[3 * n] - fb[3 * xm]) + (fb[3 * n + 1] - fb[3 * ym + 1]) + (fb[3 * n + 2] - fb[3 * zm + 2])) / g.dx;
            }
            var("cellDivB", dv.data(), 1, "float", 8);
        }
        if (!log_rhoM.p) log_rhoM.alloc(nc);
        if (want.count("cellNPles")) {
            // calcCellNPles (:355-373): macroparticles of all species per cell
            if (sp.empty()) fill(log_rhoM.p, 0.0, log_rhoM.n);
            for (size_t s = 0; s < sp.size(); ++s)
                launch("vlsv_count", k_cell_mass<Real>, cdiv(g.nc, 256), 256, 0, g, (const uint32_t*)sp[s]->sub.p,
                       sp[s]->A(), 1.0, log_rhoM.p, s == 0 ? 1 : 0);
            h.assign(nc, 0.0);
            RH_HIP(hipMemcpyAsync(h.data(), log_rhoM.p, nc * sizeof(double), hipMemcpyDeviceToHost, stream));
            RH_HIP(hipStreamSynchronize(stream));
            var("cellNPles", h.data(), 1, "float", 8);
        }
        if (want.count("n") || want.count("v") || want.count("T")) {
            // calcCellParticleBulkParameters (:376-447) per output population,
            // in order of first appearance
            std::vector<std::string> order;
            for (size_t s = 0; s < sp.size(); ++s) {
                if (!pops[s]) throw ArgError("null population name");
                if (std::find(order.begin(), order.end(), std::string(pops[s])) == order.end()) order.push_back(pops[s]);
            }
            DevBuf<double> acc, tacc;
            acc.alloc(nc * 4);
            tacc.alloc(nc);
            for (const auto& name : order) {
                std::vector<size_t> members;
                for (size_t s = 0; s < sp.size(); ++s)
                    if (name == pops[s]) members.push_back(s);
                for (int pass = 0; pass < 2; ++pass)
                    for (size_t q = 0; q < members.size(); ++q) {
                        Species& S_ = *sp[members[q]];
                        launch("vlsv_bulk", k_cell_bulk<Real>, cdiv(g.nc, 256), 256, 0, g, (const uint32_t*)S_.sub.p,
                               S_.A(), S_.d.w, pass, acc.p, tacc.p, q == 0 ? 1 : 0);
                    }
                std::vector<double> ha(nc * 4), ht(nc);
                RH_HIP(hipMemcpyAsync(ha.data(), acc.p, ha.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
                RH_HIP(hipMemcpyAsync(ht.data(), tacc.p, ht.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
                RH_HIP(hipStreamSynchronize(stream));
                const double m = sp[members[0]]->d.m, dV = g.dx * g.dx * g.dx;
                std::vector<double> dens(nc), temp(nc), vel(nc * 3);
                for (uint64_t n = 0; n < nc; ++n) {
                    const double s0 = ha[4 * n];
                    for (int l = 0; l < 3; ++l) vel[3 * n + l] = s0 > 0.0 ? ha[4 * n + 1 + l] / s0 : 0.0;
                    temp[n] = s0 > 0.0 ? ht[n] * (m / (3 * s0 * LOG_KB)) : 0.0;
                    dens[n] = s0 / dV;
                }
                if (want.count("n")) var("n_" + name, dens.data(), 1, "float", 8);
                if (want.count("T")) var("T_" + name, temp.data(), 1, "float", 8);
                if (want.count("v")) var("v_" + name, vel.data(), 3, "float", 8);
            }
        }
        if (ave_on) write_averages(want, pops, var);
        W.close();
    }

    // The temporal averages of UserDataOP::writeData (src/operator_userdata.
    // cpp:156-232): cellBAverage = sum / count; per output population
    // n_<pop>_ave = sum of the members' charge / (q dV) / count (q of the
    // first member) and v_<pop>_ave = sum of current / sum of charge; the
    // sums that are written restart, and so does the count.
    template <typename VarFn>
    void write_averages(const std::map<std::string, bool>& want, const char* const* pops, VarFn& var) {
        const uint64_t nc = (uint64_t)g.nc;
        const double inv = ave_count > 0 ? 1.0 / (double)ave_count : 0.0;
        if (want.count("cellBAverage")) {
            std::vector<double> h(nc * 3), out(nc * 3);
            RH_HIP(hipMemcpyAsync(h.data(), aveB.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
            RH_HIP(hipStreamSynchronize(stream));
            for (uint64_t n = 0; n < nc; ++n)
                for (int l = 0; l < 3; ++l) out[3 * n + l] = h[(size_t)l * nc + n] * inv;
            var("cellBAverage", out.data(), 3, "float", 8);
            fill(aveB.p, 0.0, aveB.n);
        }
        if (want.count("n_ave") || want.count("v_ave")) {
            std::vector<std::string> order;
            for (size_t s = 0; s < sp.size(); ++s) {
                if (!pops[s]) throw ArgError("null population name");
                if (std::find(order.begin(), order.end(), std::string(pops[s])) == order.end()) order.push_back(pops[s]);
            }
            const double dV = g.dx * g.dx * g.dx;
            for (const auto& name : order) {
                std::vector<double> nsum(nc, 0.0), vsum(nc * 3, 0.0), hn(nc), hv(nc * 3);
                double q0 = 0.0;
                for (size_t s = 0; s < sp.size(); ++s) {
                    if (name != pops[s]) continue;
                    if (q0 == 0.0) q0 = sp[s]->d.q;
                    RH_HIP(hipMemcpyAsync(hn.data(), sp[s]->ave_n.p, nc * sizeof(double), hipMemcpyDeviceToHost, stream));
                    RH_HIP(hipMemcpyAsync(hv.data(), sp[s]->ave_v.p, hv.size() * sizeof(double), hipMemcpyDeviceToHost,
                                          stream));
                    RH_HIP(hipStreamSynchronize(stream));
                    for (uint64_t n = 0; n < nc; ++n) {
                        nsum[n] += hn[n];
                        for (int l = 0; l < 3; ++l) vsum[3 * n + l] += hv[(size_t)l * nc + n];
                    }
                    fill(sp[s]->ave_n.p, 0.0, sp[s]->ave_n.n);
                    fill(sp[s]->ave_v.p, 0.0, sp[s]->ave_v.n);
                }
                std::vector<double> dens(nc), vel(nc * 3);
                for (uint64_t n = 0; n < nc; ++n) {
                    dens[n] = nsum[n] / (q0 * dV) * inv;
                    for (int l = 0; l < 3; ++l) vel[3 * n + l] = nsum[n] > 0 ? vsum[3 * n + l] / nsum[n] : 0.0;
                }
                if (want.count("n_ave")) var("n_" + name + "_ave", dens.data(), 1, "float", 8);
                if (want.count("v_ave")) var("v_" + name + "_ave", vel.data(), 3, "float", 8);
            }
        }
        ave_count = 0;
        RH_HIP(hipStreamSynchronize(stream));
    }

    int64_t injected(int s) override {
        Species& S_ = S(s);
        if (!S_.inj_total.p) return 0;
        unsigned long long h = 0;
        RH_HIP(hipMemcpyAsync(&h, S_.inj_total.p, sizeof(h), hipMemcpyDeviceToHost, stream));
        RH_HIP(hipStreamSynchronize(stream));
        return (int64_t)h;
    }

    // t[DIAG_Q]: count, sum vx, vy, vz, sum |v|, sum v^2, max |v| (unweighted)
    void diag_raw(int s, double (&t)[DIAG_Q]) {
        Species& S_ = S(s);
        if (S_.pushed) throw StateError("diagnostics between propagate and wait_particle_sends");
        const uint32_t* extra = S_.inj_pending ? (const uint32_t*)(S_.inj_off.p + S_.inj_ncell) : nullptr;
        launch("diag", k_diag<Real>, DIAG_BLOCKS, 256, 0, S_.A(), (const uint32_t*)(S_.sub.p + g.nkeys), extra,
               (uint32_t)S_.cap, diag_partial.p);
        std::vector<double> h(DIAG_BLOCKS * DIAG_Q);
        RH_HIP(hipMemcpyAsync(h.data(), diag_partial.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, stream));
        RH_HIP(hipStreamSynchronize(stream));
        for (int q = 0; q < DIAG_Q; ++q) t[q] = 0;
        for (int b = 0; b < DIAG_BLOCKS; ++b) {
            for (int q = 0; q < DIAG_Q - 1; ++q) t[q] += h[b * DIAG_Q + q];
            t[DIAG_Q - 1] = std::max(t[DIAG_Q - 1], h[b * DIAG_Q + DIAG_Q - 1]);
        }
    }

    void diagnostics(int s, double* out) override {
        double t[DIAG_Q];
        diag_raw(s, t);
        const double w = S(s).d.w;
        out[0] = t[0] * w; out[1] = t[1] * w; out[2] = t[2] * w; out[3] = t[3] * w; out[4] = t[5] * w;
    }

    // counter totals: the context's own and every species' push counters
    void counter_totals(unsigned long long* h) {
        std::vector<unsigned long long> hs((size_t)C_N * COUNTER_SLOTS);
        for (int q = 0; q < C_N; ++q) h[q] = 0;
        for (size_t b = 0; b <= sp.size(); ++b) {
            const unsigned long long* src = b == 0 ? dcount.p : sp[b - 1]->pcount.p;
            RH_HIP(hipMemcpyAsync(hs.data(), src, sizeof(unsigned long long) * hs.size(), hipMemcpyDeviceToHost, stream));
            RH_HIP(hipStreamSynchronize(stream));
            for (int sl = 0; sl < COUNTER_SLOTS; ++sl)
                for (int q = 0; q < C_N; ++q) h[q] += hs[(size_t)sl * C_N + q];
        }
    }

    unsigned long long species_counter(Species& S_, int id) {
        std::vector<unsigned long long> hs((size_t)C_N * COUNTER_SLOTS);
        RH_HIP(hipMemcpyAsync(hs.data(), S_.pcount.p, sizeof(unsigned long long) * hs.size(), hipMemcpyDeviceToHost, stream));
        RH_HIP(hipStreamSynchronize(stream));
        unsigned long long h = 0;
        for (int sl = 0; sl < COUNTER_SLOTS; ++sl) h += hs[(size_t)sl * C_N + id];
        return h;
    }

    void counters(double* out) override {
        unsigned long long h[C_N];
        counter_totals(h);
        for (int q = 0; q < RH_N_COUNTERS; ++q) out[q] = (double)h[q];
        if (h[C_RANK_OVERFLOW])
            throw CapError("particle sort overflow (a subcell holds more than 65535 particles of one species, "
                           "or the particle store is full)");
    }

    void synchronize() override { RH_HIP(hipStreamSynchronize(stream)); }

    void set_timing(bool on) override {
        if (!on && timing) resolve_timers();
        timing = on;
    }

    void set_timing_only(const char* name) override { timing_only = name ? name : ""; }
    void set_species_streams(bool on) override { species_streams = on; }

    void kernel_time(const char* name, double* ms, int64_t* launches) override {
        resolve_timers();
        auto it = timers.find(name ? name : "");
        if (ms) *ms = it == timers.end() ? 0.0 : it->second.ms;
        if (launches) *launches = it == timers.end() ? 0 : it->second.launches;
    }

    void reset_timing() override {
        resolve_timers();
        timers.clear();
    }

    void layout(int64_t* out) override {
        out[0] = g.ntiles; out[1] = g.tx; out[2] = g.ty; out[3] = g.tz; out[4] = g.nc;
        out[5] = (int64_t)(6 * sizeof(Real) + sizeof(uint32_t));
    }
};

}  // namespace rh

struct rh_ctx {
    std::unique_ptr<rh::Ctx> impl;
    std::string err;
};

template <typename F>
static int guarded(rh_ctx* ctx, F&& f) {
    if (!ctx || !ctx->impl) return RH_ERR_ARG;
    try {
        f(*ctx->impl);
        return RH_OK;
    } catch (const rh::ArgError& e) {
        ctx->err = e.what();
        return RH_ERR_ARG;
    } catch (const rh::CapError& e) {
        ctx->err = e.what();
        return RH_ERR_CAPACITY;
    } catch (const rh::StateError& e) {
        ctx->err = e.what();
        return RH_ERR_STATE;
    } catch (const rh::HipError& e) {
        ctx->err = e.what();
        return RH_ERR_HIP;
    } catch (const std::exception& e) {
        ctx->err = e.what();
        return RH_ERR_HIP;
    }
}

extern "C" {

int rh_create(const rh_grid_desc* grid, const rh_params* params, int32_t precision, rh_ctx** out) {
    if (!grid || !params || !out) return RH_ERR_ARG;
    *out = nullptr;
    auto* c = new rh_ctx();
    try {
        if (precision == RH_FP32) c->impl.reset(new rh::Sim<float>(grid, params));
        else if (precision == RH_FP64) c->impl.reset(new rh::Sim<double>(grid, params));
        else throw rh::ArgError("precision must be RH_FP32 or RH_FP64");
    } catch (const rh::ArgError& e) {
        delete c;
        return RH_ERR_ARG;
    } catch (const std::exception& e) {
        int dev = 0;
        const bool nodev = hipGetDeviceCount(&dev) != hipSuccess || dev == 0;
        delete c;
        return nodev ? RH_ERR_NO_DEVICE : RH_ERR_HIP;
    }
    *out = c;
    return RH_OK;
}

int rh_destroy(rh_ctx* ctx) {
    if (!ctx) return RH_ERR_ARG;
    delete ctx;
    return RH_OK;
}

const char* rh_last_error(const rh_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int rh_add_species(rh_ctx* ctx, const rh_species_desc* desc, int32_t* id) {
    if (!desc || !id) return RH_ERR_ARG;
    return guarded(ctx, [&](rh::Ctx& c) { *id = c.add_species(desc); });
}

int rh_upload_field(rh_ctx* ctx, int32_t f, const double* d) {
    return guarded(ctx, [&](rh::Ctx& c) { c.upload_field(f, d); });
}
int rh_download_field(rh_ctx* ctx, int32_t f, double* d) {
    return guarded(ctx, [&](rh::Ctx& c) { c.download_field(f, d); });
}
int rh_upload_flags(rh_ctx* ctx, int32_t f, const uint8_t* d) {
    return guarded(ctx, [&](rh::Ctx& c) { c.upload_flags(f, d); });
}
int rh_upload_particles(rh_ctx* ctx, int32_t s, int64_t n, const int64_t* cell, const double* st) {
    return guarded(ctx, [&](rh::Ctx& c) { c.upload_particles(s, n, cell, st); });
}
int rh_num_particles(rh_ctx* ctx, int32_t s, int64_t* n) {
    if (!n) return RH_ERR_ARG;
    return guarded(ctx, [&](rh::Ctx& c) { *n = c.num_particles(s); });
}
int rh_download_particles(rh_ctx* ctx, int32_t s, int64_t cap, int64_t* cell, double* st) {
    return guarded(ctx, [&](rh::Ctx& c) { c.download_particles(s, cap, cell, st); });
}
int rh_setup_get_fields(rh_ctx* ctx) {
    return guarded(ctx, [&](rh::Ctx&) {});
}
int rh_propagate(rh_ctx* ctx, int32_t s) {
    return guarded(ctx, [&](rh::Ctx& c) { c.propagate(s); });
}
int rh_wait_particle_sends(rh_ctx* ctx, int32_t s) {
    return guarded(ctx, [&](rh::Ctx& c) { c.wait_particle_sends(s); });
}
int rh_clear_accumulation_arrays(rh_ctx* ctx) {
    return guarded(ctx, [&](rh::Ctx& c) { c.clear_accumulation(); });
}
int rh_accumulate(rh_ctx* ctx, int32_t s) {
    return guarded(ctx, [&](rh::Ctx& c) { c.accumulate(s); });
}
int rh_add_remote_updates(rh_ctx* ctx) {
    return guarded(ctx, [&](rh::Ctx& c) { c.add_remote_updates(); });
}
int rh_apply_boundary_conditions(rh_ctx* ctx, int32_t s) {
    return guarded(ctx, [&](rh::Ctx& c) { c.apply_bcs(s); });
}
int rh_propagate_B(rh_ctx* ctx) {
    return guarded(ctx, [&](rh::Ctx& c) { c.propagate_B(); });
}
int rh_step(rh_ctx* ctx, int32_t n) {
    if (n < 0) return RH_ERR_ARG;
    return guarded(ctx, [&](rh::Ctx& c) { c.step(n); });
}
int rh_diagnostics(rh_ctx* ctx, int32_t s, double out[5]) {
    if (!out) return RH_ERR_ARG;
    return guarded(ctx, [&](rh::Ctx& c) { c.diagnostics(s, out); });
}
int rh_counters(rh_ctx* ctx, double out[RH_N_COUNTERS]) {
    if (!out) return RH_ERR_ARG;
    return guarded(ctx, [&](rh::Ctx& c) { c.counters(out); });
}
int rh_synchronize(rh_ctx* ctx) {
    return guarded(ctx, [&](rh::Ctx& c) { c.synchronize(); });
}
int rh_set_timing(rh_ctx* ctx, int32_t on) {
    return guarded(ctx, [&](rh::Ctx& c) { c.set_timing(on != 0); });
}
int rh_set_timing_only(rh_ctx* ctx, const char* kernel) {
    return guarded(ctx, [&](rh::Ctx& c) { c.set_timing_only(kernel); });
}

int rh_set_species_streams(rh_ctx* ctx, int32_t enable) {
    return guarded(ctx, [&](rh::Ctx& c) { c.set_species_streams(enable != 0); });
}
int rh_kernel_times(rh_ctx* ctx, const char* kernel, double* ms, int64_t* launches) {
    return guarded(ctx, [&](rh::Ctx& c) { c.kernel_time(kernel, ms, launches); });
}
int rh_reset_timing(rh_ctx* ctx) {
    return guarded(ctx, [&](rh::Ctx& c) { c.reset_timing(); });
}
int rh_set_denoiser(rh_ctx* ctx, rh_denoiser* dn, double max_rel_change) {
    return guarded(ctx, [&](rh::Ctx& c) { c.set_denoiser(dn, max_rel_change); });
}
int rh_denoise_moments(rh_ctx* ctx, rh_denoiser* dn, double max_rel_change) {
    return guarded(ctx, [&](rh::Ctx& c) { c.denoise(dn, max_rel_change); });
}
int rh_set_injector(rh_ctx* ctx, int32_t s, const rh_injector_desc* desc) {
    return guarded(ctx, [&](rh::Ctx& c) { c.set_injector(s, desc); });
}
int rh_inject_particles(rh_ctx* ctx, int32_t s) {
    return guarded(ctx, [&](rh::Ctx& c) { c.inject(s); });
}
int rh_set_timestep(rh_ctx* ctx, int64_t t) {
    return guarded(ctx, [&](rh::Ctx& c) { c.set_timestep(t); });
}
int rh_get_timestep(rh_ctx* ctx, int64_t* t) {
    if (!t) return RH_ERR_ARG;
    return guarded(ctx, [&](rh::Ctx& c) { *t = c.timestep(); });
}
int rh_injected(rh_ctx* ctx, int32_t s, int64_t* n) {
    if (!n) return RH_ERR_ARG;
    return guarded(ctx, [&](rh::Ctx& c) { *n = c.injected(s); });
}
int rh_log_begin(rh_ctx* ctx, double t) {
    return guarded(ctx, [&](rh::Ctx& c) { c.log_begin(t); });
}
int rh_log_values(rh_ctx* ctx, double t, int32_t include_inner_cells, double field[41], double* pop) {
    if (!field || !pop) return RH_ERR_ARG;
    return guarded(ctx, [&](rh::Ctx& c) { c.log_values(t, include_inner_cells, field, pop); });
}
int rh_log_open(rh_ctx* ctx, const char* directory, const char* const* species_names, int32_t precision,
                int32_t include_inner_cells, double max_B_stop) {
    return guarded(ctx, [&](rh::Ctx& c) { c.log_open(directory, species_names, precision, include_inner_cells, max_B_stop); });
}
int rh_log_write(rh_ctx* ctx, double t) {
    return guarded(ctx, [&](rh::Ctx& c) { c.log_write(t); });
}
int rh_write_vlsv(rh_ctx* ctx, const char* filename, double t, int64_t timestep, const char* const* variables,
                  int32_t nvariables, const char* const* population_names) {
    if (nvariables < 0) return RH_ERR_ARG;
    return guarded(ctx, [&](rh::Ctx& c) { c.write_vlsv(filename, t, timestep, variables, nvariables, population_names); });
}
int rh_set_temporal_averages(rh_ctx* ctx, int32_t enable) {
    return guarded(ctx, [&](rh::Ctx& c) { c.set_temporal_averages(enable != 0); });
}
int rh_layout_info(rh_ctx* ctx, int64_t out[6]) {
    if (!out) return RH_ERR_ARG;
    return guarded(ctx, [&](rh::Ctx& c) { c.layout(out); });
}

}  // extern "C"
