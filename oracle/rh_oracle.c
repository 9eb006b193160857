/*
 * rh_oracle.c -- FP64 CPU restatement of RHybrid's per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; also the CPU
 * baseline that bench.py times).  Every function names the reference lines it
 * restates.  The reference cannot be compiled here (it needs the un-vendored
 * Corsair/ParGrid/VLSV stack), so this file is written from reading it.
 *
 * Conventions shared with the reference:
 *  - cell n = i + nx*(j + ny*k); vector arrays interleaved n*3+l;
 *  - faceB[n*3+c] is B_c on the +c face of cell n; node n is the upper
 *    (+x,+y,+z) corner of cell n (src/user.cpp:2111-2141, doc/cell_indexing.pdf);
 *  - particles are per cell with cell-local coordinates in [0,dx).
 * The 3x3x3 gather buffer `aa` of fetchData (src/hybrid_propagator.cpp:1894)
 * is indexed here as AI(a,b,c) = a + 3b + 9c with a,b,c in {0,1,2} meaning the
 * neighbour at offset (a-1,b-1,c-1), which is what block::arrayIndex means
 * for Corsair's 1x1x1 blocks (src/user.cpp:164-167).
 */
#include "rh_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* CODATA values; Corsair's constants:: table is not in the reference tree
 * (tools/trace_test_particles.py:47-48 carries the same e and m_p). */
#define ORC_E    1.60217653e-19
#define ORC_KB   1.3806505e-23
#define ORC_MU0  (4.0e-7 * M_PI)

#define MAX_SPECIES 8
#define AI(a, b, c) ((a) + 3 * (b) + 9 * (c))

struct orc_sim {
    orc_grid g;
    orc_params p;
    int nc;
    double* f[ORC_N_FIELDS];
    uint8_t* flag[ORC_N_FLAGS];
    uint32_t* nflags;     /* bit AI(a,b,c) set if that neighbour exists */
    int* exterior;        /* cells missing any neighbour */
    int nexterior;
    int ext_reversed;
    int nspecies;
    orc_species sp[MAX_SPECIES];
    int64_t np[MAX_SPECIES];
    double* pst[MAX_SPECIES];     /* AoS state[7] per particle */
    int64_t* cstart[MAX_SPECIES]; /* nc+1 offsets */
    double counters[ORC_N_COUNTERS];
    double* accbuf;               /* per cell 27 x (rho, Jx, Jy, Jz) */
    int nthreads;
    int64_t timestep;             /* Simulation::timestep */
    struct orc_inj {
        int kind;
        uint32_t key[2];
        double vth;
        int axis, pos_wall;       /* solar wind: normal axis; 1 = the +axis wall */
        double rate;              /* solar wind: macroparticles per cell per dt */
        double x0, xm, c, d, cxm; /* sampleDerivGaussian constants for x0 = U/vth */
        double* cell_rate;        /* exosphere */
        double R2_exobase;
        double U[3], box[6];      /* uniform: bulk velocity, filled box of cell centres */
        int64_t injected;
    } inj[MAX_SPECIES];
    /* log counters per species (Hybrid::logCounterParticle*): escape, impact
     * and inject sums of w and w*v^2, injected macroparticles, maxVi clamps */
    double esc_w[MAX_SPECIES], esc_wv2[MAX_SPECIES], imp_w[MAX_SPECIES], imp_wv2[MAX_SPECIES];
    double inj_w[MAX_SPECIES], inj_wv2[MAX_SPECIES], inj_macro[MAX_SPECIES], maxvi[MAX_SPECIES];
    /* counter values and time at the last log line */
    double log_t0;
    double log_cnt0[ORC_N_COUNTERS];
};

static const int FIELD_DIM[ORC_N_FIELDS] = {3, 3, 3, 3, 3, 3, 1, 1, 3, 3, 3, 3, 3, 1};

/* neighbour-existence masks, cf. Hybrid::X_POS_EXISTS etc. (src/user.cpp:176-181) */
#define X_POS (1u << AI(2, 1, 1))
#define X_NEG (1u << AI(0, 1, 1))
#define Y_POS (1u << AI(1, 2, 1))
#define Y_NEG (1u << AI(1, 0, 1))
#define Z_POS (1u << AI(1, 1, 2))
#define Z_NEG (1u << AI(1, 1, 0))
#define ALL_NBRS 0x7FFFFFFu

static inline void decompose(const orc_grid* g, int n, int* i, int* j, int* k) {
    *i = n % g->nx;
    *j = (n / g->nx) % g->ny;
    *k = n / (g->nx * g->ny);
}

/* Neighbour of cell n at offset (a,b,c); -1 if it does not exist
 * (ParGrid neighbour topology: periodic wrap or a missing neighbour). */
static int nbr(const orc_grid* g, int n, int a, int b, int c) {
    int i, j, k;
    decompose(g, n, &i, &j, &k);
    i += a; j += b; k += c;
    if (i < 0 || i >= g->nx) { if (!g->periodic[0]) return -1; i = (i % g->nx + g->nx) % g->nx; }
    if (j < 0 || j >= g->ny) { if (!g->periodic[1]) return -1; j = (j % g->ny + g->ny) % g->ny; }
    if (k < 0 || k >= g->nz) { if (!g->periodic[2]) return -1; k = (k % g->nz + g->nz) % g->nz; }
    return i + g->nx * (j + g->ny * k);
}

/* fetchData (src/hybrid_propagator.cpp:1894-2094): gather the 3x3x3
 * neighbourhood; entries of missing neighbours are left untouched. */
static void fetch(const orc_sim* s, const double* D, double* aa, int n, int dim) {
    for (int c = 0; c < 3; ++c)
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) {
                int m = nbr(&s->g, n, a - 1, b - 1, c - 1);
                if (m < 0) continue;
                for (int l = 0; l < dim; ++l) aa[AI(a, b, c) * dim + l] = D[(size_t)m * dim + l];
            }
}

static void fill_nan(double* aa, int count) {
    for (int i = 0; i < count; ++i) aa[i] = NAN;
}

static inline double sqr(double x) { return x * x; }

/* ------------------------------------------------------------------ setup */

orc_sim* orc_create(const orc_grid* g, const orc_params* p) {
    if (g->nx < 3 || g->ny < 3 || g->nz < 3 || g->dx <= 0) return NULL;
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    s->g = *g;
    s->p = *p;
    s->nc = g->nx * g->ny * g->nz;
    for (int f = 0; f < ORC_N_FIELDS; ++f) s->f[f] = (double*)calloc((size_t)s->nc * FIELD_DIM[f], sizeof(double));
    for (int f = 0; f < ORC_N_FLAGS; ++f) s->flag[f] = (uint8_t*)calloc((size_t)s->nc, 1);
    s->nflags = (uint32_t*)malloc(sizeof(uint32_t) * s->nc);
    s->exterior = (int*)malloc(sizeof(int) * s->nc);
    for (int n = 0; n < s->nc; ++n) {
        uint32_t m = 0;
        for (int c = 0; c < 3; ++c)
            for (int b = 0; b < 3; ++b)
                for (int a = 0; a < 3; ++a)
                    if (nbr(g, n, a - 1, b - 1, c - 1) >= 0) m |= 1u << AI(a, b, c);
        s->nflags[n] = m;
        if (m != ALL_NBRS) s->exterior[s->nexterior++] = n;
    }
    s->accbuf = (double*)malloc(sizeof(double) * (size_t)s->nc * 108);
    s->nthreads = 0;
    return s;
}

void orc_destroy(orc_sim* s) {
    if (!s) return;
    for (int f = 0; f < ORC_N_FIELDS; ++f) free(s->f[f]);
    for (int f = 0; f < ORC_N_FLAGS; ++f) free(s->flag[f]);
    for (int i = 0; i < s->nspecies; ++i) { free(s->pst[i]); free(s->cstart[i]); free(s->inj[i].cell_rate); }
    free(s->nflags); free(s->exterior); free(s->accbuf);
    free(s);
}

int orc_add_species(orc_sim* s, const orc_species* sp) {
    if (s->nspecies >= MAX_SPECIES) return -1;
    int id = s->nspecies++;
    s->sp[id] = *sp;
    s->np[id] = 0;
    s->pst[id] = NULL;
    s->cstart[id] = (int64_t*)calloc((size_t)s->nc + 1, sizeof(int64_t));
    return id;
}

int orc_num_cells(const orc_sim* s) { return s->nc; }
int orc_field_dim(int id) { return (id >= 0 && id < ORC_N_FIELDS) ? FIELD_DIM[id] : 0; }
double* orc_field(orc_sim* s, int id) { return (id >= 0 && id < ORC_N_FIELDS) ? s->f[id] : NULL; }
uint8_t* orc_flag(orc_sim* s, int id) { return (id >= 0 && id < ORC_N_FLAGS) ? s->flag[id] : NULL; }
void orc_set_threads(orc_sim* s, int nthreads) {
    s->nthreads = nthreads;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
}
void orc_set_exterior_order_reversed(orc_sim* s, int r) { s->ext_reversed = r; }
double orc_counter(const orc_sim* s, int id) { return (id >= 0 && id < ORC_N_COUNTERS) ? s->counters[id] : 0.0; }
void orc_reset_counters(orc_sim* s) { memset(s->counters, 0, sizeof(s->counters)); }

/* Stable counting sort of (cell, state) records into the CSR store. */
int orc_set_particles(orc_sim* s, int sp, int64_t n, const int64_t* cell, const double* st) {
    if (sp < 0 || sp >= s->nspecies) return -1;
    int64_t* cs = s->cstart[sp];
    memset(cs, 0, sizeof(int64_t) * ((size_t)s->nc + 1));
    for (int64_t p = 0; p < n; ++p) {
        if (cell[p] < 0 || cell[p] >= s->nc) return -2;
        cs[cell[p] + 1]++;
    }
    for (int c = 0; c < s->nc; ++c) cs[c + 1] += cs[c];
    free(s->pst[sp]);
    s->pst[sp] = (double*)malloc(sizeof(double) * 7 * (size_t)(n > 0 ? n : 1));
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * s->nc);
    memcpy(fill, cs, sizeof(int64_t) * s->nc);
    for (int64_t p = 0; p < n; ++p) {
        int64_t d = fill[cell[p]]++;
        memcpy(s->pst[sp] + 7 * d, st + 7 * p, 7 * sizeof(double));
    }
    free(fill);
    s->np[sp] = n;
    return 0;
}

int64_t orc_num_particles(const orc_sim* s, int sp) { return (sp >= 0 && sp < s->nspecies) ? s->np[sp] : -1; }
const int64_t* orc_cell_start(const orc_sim* s, int sp) { return s->cstart[sp]; }

int orc_get_particles(const orc_sim* s, int sp, int64_t* cell, double* st) {
    if (sp < 0 || sp >= s->nspecies) return -1;
    for (int c = 0; c < s->nc; ++c)
        for (int64_t p = s->cstart[sp][c]; p < s->cstart[sp][c + 1]; ++p) {
            if (cell) cell[p] = c;
            if (st) memcpy(st + 7 * p, s->pst[sp] + 7 * p, 7 * sizeof(double));
        }
    return 0;
}

/* ------------------------------------------------------ field gather */

/* face2r (src/hybrid_propagator.cpp:1740-1778): linear interpolation of each
 * face component along its own axis only. */
static void face2r(const orc_sim* s, const double* r, int64_t n, double* out) {
    double aa[81];
    fill_nan(aa, 81);
    fetch(s, s->f[ORC_FACE_B], aa, (int)n, 3);
    const double dx = s->g.dx;
    const double x = r[0] / dx, y = r[1] / dx, z = r[2] / dx;
    const int i = (int)floor(x), j = (int)floor(y), k = (int)floor(z);
    const double w111x = x - i, w111y = y - j, w111z = z - k;
    const double w011x = (i + 1) - x, w101y = (j + 1) - y, w110z = (k + 1) - z;
    out[0] = w111x * aa[AI(i + 1, j + 1, k + 1) * 3 + 0] + w011x * aa[AI(i + 0, j + 1, k + 1) * 3 + 0];
    out[1] = w111y * aa[AI(i + 1, j + 1, k + 1) * 3 + 1] + w101y * aa[AI(i + 1, j + 0, k + 1) * 3 + 1];
    out[2] = w111z * aa[AI(i + 1, j + 1, k + 1) * 3 + 2] + w110z * aa[AI(i + 1, j + 1, k + 0) * 3 + 2];
}

/* cell2r (src/hybrid_propagator.cpp:1838-1858): nearest grid point. */
static void cell2r(const orc_sim* s, const double* D, int64_t n, double* out) {
    for (int l = 0; l < 3; ++l) out[l] = D[n * 3 + l];
}

/* getFields (src/hybrid_propagator.cpp:1879-1891). */
void orc_get_fields(orc_sim* s, int64_t n, const double r[3], double B[3], double Ue[3], double Ep[3]) {
    face2r(s, r, n, B);
    cell2r(s, s->f[ORC_CELL_UE], n, Ue);
    if (s->p.use_e_pressure) cell2r(s, s->f[ORC_CELL_EP], n, Ep);
    else Ep[0] = Ep[1] = Ep[2] = 0.0;
}

/* ------------------------------------------------------------ push */

/* BorisBuneman::propagate (src/particle_propagator_boris_buneman.h:64-194),
 * called per cell by propagateCell (:196-207). */
void orc_push(orc_sim* s, int sp) {
    const orc_species* S = &s->sp[sp];
    const double dt = s->p.dt;
    const double qmideltT2 = 0.5 * S->q * dt / S->m;
    double nvi = 0.0;
    const double* crd0 = &s->g.x0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : nvi)
    for (int n = 0; n < s->nc; ++n) {
        int accelerate = S->accelerate;
        if (s->nflags[n] != ALL_NBRS || s->p.initial_flow_through) accelerate = 0;
        int ci, cj, ck;
        decompose(&s->g, n, &ci, &cj, &ck);
        const double xBlock = crd0[0] + ci * s->g.dx;
        const double yBlock = crd0[1] + cj * s->g.dx;
        const double zBlock = crd0[2] + ck * s->g.dx;
        for (int64_t p = s->cstart[sp][n]; p < s->cstart[sp][n + 1]; ++p) {
            double* st = s->pst[sp] + 7 * p;
            if (accelerate) {
                double E[3], B[3], Ue[3], Ep[3];
                const double r[3] = {st[0], st[1], st[2]};
                const double rG[3] = {r[0] + xBlock, r[1] + yBlock, r[2] + zBlock};
                orc_get_fields(s, n, r, B, Ue, Ep);
                /* E = B x Ue = -Ue x B */
                E[0] = B[1] * Ue[2] - B[2] * Ue[1];
                E[1] = B[2] * Ue[0] - B[0] * Ue[2];
                E[2] = B[0] * Ue[1] - B[1] * Ue[0];
                if (s->p.use_e_pressure) for (int i = 0; i < 3; ++i) E[i] += Ep[i];
                const double dvx = qmideltT2 * E[0], dvy = qmideltT2 * E[1], dvz = qmideltT2 * E[2];
                const double tx = qmideltT2 * B[0], ty = qmideltT2 * B[1], tz = qmideltT2 * B[2];
                const double t2 = tx * tx + ty * ty + tz * tz;
                const double b2 = 2. / (1. + t2);
                const double sx = b2 * tx, sy = b2 * ty, sz = b2 * tz;
                const double vmx = st[3] + dvx, vmy = st[4] + dvy, vmz = st[5] + dvz;
                const double v0x = vmx + vmy * tz - vmz * ty;
                const double v0y = vmy + vmz * tx - vmx * tz;
                const double v0z = vmz + vmx * ty - vmy * tx;
                const double vpx = vmx + v0y * sz - v0z * sy;
                const double vpy = vmy + v0z * sx - v0x * sz;
                const double vpz = vmz + v0x * sy - v0y * sx;
                st[3] = vpx + dvx;
                st[4] = vpy + dvy;
                st[5] = vpz + dvz;
                if (s->p.use_gravity) {
                    const double r3 = pow(sqrt(sqr(rG[0]) + sqr(rG[1]) + sqr(rG[2])), 3);
                    if (r3 > 0) {
                        const double sg = -s->p.GMdt / r3;
                        st[3] += sg * rG[0]; st[4] += sg * rG[1]; st[5] += sg * rG[2];
                    }
                }
                const double v2 = sqr(st[3]) + sqr(st[4]) + sqr(st[5]);
                if (v2 > s->p.max_ion_speed2) {
                    const double norm = sqrt(s->p.max_ion_speed2 / v2);
                    st[3] *= norm; st[4] *= norm; st[5] *= norm;
                    nvi += 1.0;
                }
            }
            st[0] += dt * st[3];
            st[1] += dt * st[4];
            st[2] += dt * st[5];
        }
    }
    s->counters[ORC_CNT_MAX_VI] += nvi;
    s->maxvi[sp] += nvi;
}

/* -------------------------------------------------------- migration */

/* Corsair's ParticleListSkeleton re-homes particles whose cell-local
 * coordinates left [0,dx) (external, not in the reference tree).  Restated:
 * the destination is the neighbour at offset floor(x/dx) per axis, the local
 * coordinate is shifted by the same number of cells, particles whose
 * destination does not exist (outside a non-periodic domain) or is more than
 * one cell away are dropped.  The result is stored as a stable counting sort
 * by destination cell (arrivals in ascending source-cell order). */
void orc_migrate(orc_sim* s, int sp) {
    const int64_t np = s->np[sp];
    const double dx = s->g.dx;
    const int nc = s->nc;
    int64_t* dest = (int64_t*)malloc(sizeof(int64_t) * (size_t)(np > 0 ? np : 1));
    int64_t* cnt = (int64_t*)calloc((size_t)nc * 27, sizeof(int64_t));
    double* st = s->pst[sp];
    const int64_t* cs = s->cstart[sp];
#pragma omp parallel for schedule(dynamic, 64)
    for (int n = 0; n < nc; ++n) {
        for (int64_t p = cs[n]; p < cs[n + 1]; ++p) {
            double* q = st + 7 * p;
            int off[3];
            int ok = 1;
            for (int l = 0; l < 3; ++l) {
                double o = floor(q[l] / dx);
                if (o < -1.0 || o > 1.0 || q[l] != q[l]) { ok = 0; off[l] = 0; continue; }
                off[l] = (int)o;
                q[l] -= off[l] * dx;
                if (q[l] >= dx) q[l] = nextafter(dx, 0.0);
                if (q[l] < 0.0) q[l] = 0.0;
            }
            int d = ok ? AI(off[0] + 1, off[1] + 1, off[2] + 1) : -1;
            if (d >= 0 && nbr(&s->g, n, off[0], off[1], off[2]) < 0) d = -1;
            dest[p] = d;
            if (d >= 0) cnt[(size_t)n * 27 + d]++;
        }
    }
    /* arrivals at cell t, in ascending source-cell order */
    int64_t* newcs = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));
    int64_t* base = (int64_t*)malloc(sizeof(int64_t) * (size_t)nc * 27);
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nc; ++t) {
        int64_t tot = 0;
        for (int d = 0; d < 27; ++d) {
            int src = nbr(&s->g, t, 1 - d % 3, 1 - (d / 3) % 3, 1 - d / 9);
            if (src >= 0) tot += cnt[(size_t)src * 27 + d];
        }
        newcs[t + 1] = tot;
    }
    for (int t = 0; t < nc; ++t) newcs[t + 1] += newcs[t];
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nc; ++t) {
        int srcs[27], dirs[27], m = 0;
        for (int d = 0; d < 27; ++d) {
            int src = nbr(&s->g, t, 1 - d % 3, 1 - (d / 3) % 3, 1 - d / 9);
            if (src < 0) continue;
            int at = m++;
            while (at > 0 && srcs[at - 1] > src) { srcs[at] = srcs[at - 1]; dirs[at] = dirs[at - 1]; --at; }
            srcs[at] = src; dirs[at] = d;
        }
        int64_t off = newcs[t];
        for (int q = 0; q < m; ++q) {
            base[(size_t)srcs[q] * 27 + dirs[q]] = off;
            off += cnt[(size_t)srcs[q] * 27 + dirs[q]];
        }
    }
    const int64_t nnew = newcs[nc];
    double* nst = (double*)malloc(sizeof(double) * 7 * (size_t)(nnew > 0 ? nnew : 1));
#pragma omp parallel for schedule(dynamic, 64)
    for (int n = 0; n < nc; ++n) {
        int64_t run[27];
        for (int d = 0; d < 27; ++d) run[d] = base[(size_t)n * 27 + d];
        for (int64_t p = cs[n]; p < cs[n + 1]; ++p) {
            if (dest[p] < 0) continue;
            memcpy(nst + 7 * run[dest[p]]++, st + 7 * p, 7 * sizeof(double));
        }
    }
    free(s->pst[sp]);
    s->pst[sp] = nst;
    memcpy(s->cstart[sp], newcs, sizeof(int64_t) * ((size_t)nc + 1));
    s->np[sp] = nnew;
    free(dest); free(cnt); free(newcs); free(base);
}

/* ----------------------------------------------------- accumulation */

/* Accumulator::clearAccumulationArrays (src/particle_accumulator.cpp:337-362). */
void orc_clear_accumulation(orc_sim* s) {
    memset(s->f[ORC_CELL_RHOQ], 0, sizeof(double) * s->nc);
    memset(s->f[ORC_CELL_JI], 0, sizeof(double) * 3 * (size_t)s->nc);
}

/* Accumulator::accumulateCell (src/particle_accumulator.cpp:43-156) followed
 * by block::addValues3D into own + 26 neighbours.  addValues3D is Corsair's
 * (external); it adds buffer element (a,b,c) to the neighbour at offset
 * (a-1,b-1,c-1) and skips missing neighbours.  The reference scatters cell by
 * cell in ascending cell order; gathering each target's contributions in
 * ascending source-cell order adds exactly the same terms in the same order. */
void orc_accumulate(orc_sim* s, int sp) {
    const orc_species* S = &s->sp[sp];
    if (!S->accumulate) return;
    const double dx = s->g.dx;
    const double q = S->q;
    double* buf = s->accbuf;
#pragma omp parallel for schedule(dynamic, 64)
    for (int n = 0; n < s->nc; ++n) {
        double* acc = buf + (size_t)n * 108; /* acc1[27] then acc2[81] */
        double* acc1 = acc;
        double* acc2 = acc + 27;
        for (int i = 0; i < 108; ++i) acc[i] = 0.0;
        for (int64_t p = s->cstart[sp][n]; p < s->cstart[sp][n + 1]; ++p) {
            const double* st = s->pst[sp] + 7 * p;
            const double x = st[0], y = st[1], z = st[2];
            const double wq = st[6] * q;
            const double v[3] = {st[3], st[4], st[5]};
            int i = (int)floor(x / dx), j = (int)floor(y / dx), k = (int)floor(z / dx);
            if (x < (i + 0.5) * dx) --i;
            if (y < (j + 0.5) * dx) --j;
            if (z < (k + 0.5) * dx) --k;
            ++i; ++j; ++k;
            const double x0 = (i - 1 + 0.5) * dx, y0 = (j - 1 + 0.5) * dx, z0 = (k - 1 + 0.5) * dx;
            const double w_x = (x - x0) / dx, w_y = (y - y0) / dx, w_z = (z - z0) / dx;
            const double w000 = (1 - w_x) * (1 - w_y) * (1 - w_z) * wq;
            const double w100 = w_x * (1 - w_y) * (1 - w_z) * wq;
            const double w010 = (1 - w_x) * w_y * (1 - w_z) * wq;
            const double w110 = w_x * w_y * (1 - w_z) * wq;
            const double w001 = (1 - w_x) * (1 - w_y) * w_z * wq;
            const double w101 = w_x * (1 - w_y) * w_z * wq;
            const double w011 = (1 - w_x) * w_y * w_z * wq;
            const double w111 = w_x * w_y * w_z * wq;
            const int i000 = AI(i, j, k), i100 = AI(i + 1, j, k), i010 = AI(i, j + 1, k), i110 = AI(i + 1, j + 1, k);
            const int i001 = AI(i, j, k + 1), i101 = AI(i + 1, j, k + 1), i011 = AI(i, j + 1, k + 1), i111 = AI(i + 1, j + 1, k + 1);
            acc1[i000] += w000; acc1[i100] += w100; acc1[i010] += w010; acc1[i110] += w110;
            acc1[i001] += w001; acc1[i101] += w101; acc1[i011] += w011; acc1[i111] += w111;
            for (int l = 0; l < 3; ++l) {
                acc2[i000 * 3 + l] += w000 * v[l]; acc2[i100 * 3 + l] += w100 * v[l];
                acc2[i010 * 3 + l] += w010 * v[l]; acc2[i110 * 3 + l] += w110 * v[l];
                acc2[i001 * 3 + l] += w001 * v[l]; acc2[i101 * 3 + l] += w101 * v[l];
                acc2[i011 * 3 + l] += w011 * v[l]; acc2[i111 * 3 + l] += w111 * v[l];
            }
        }
    }
    double* rho = s->f[ORC_CELL_RHOQ];
    double* J = s->f[ORC_CELL_JI];
#pragma omp parallel for schedule(static)
    for (int t = 0; t < s->nc; ++t) {
        int srcs[27], slot[27], m = 0;
        for (int d = 0; d < 27; ++d) {
            /* source at offset -(d-offset) from t; t is element d of its buffer */
            int src = nbr(&s->g, t, 1 - d % 3, 1 - (d / 3) % 3, 1 - d / 9);
            if (src < 0) continue;
            int at = m++;
            while (at > 0 && srcs[at - 1] > src) { srcs[at] = srcs[at - 1]; slot[at] = slot[at - 1]; --at; }
            srcs[at] = src; slot[at] = d;
        }
        for (int q2 = 0; q2 < m; ++q2) {
            const double* acc = buf + (size_t)srcs[q2] * 108;
            const int d = slot[q2];
            rho[t] += acc[d];
            for (int l = 0; l < 3; ++l) J[(size_t)t * 3 + l] += acc[27 + d * 3 + l];
        }
    }
}

/* Accumulator::addRemoteUpdates, finalisation part
 * (src/particle_accumulator.cpp:318-333); on one process there are no
 * remote buffers to fold. */
void orc_add_remote_updates(orc_sim* s) {
    const double dV = s->g.dx * s->g.dx * s->g.dx;
    double cnt = 0;
    double* rho = s->f[ORC_CELL_RHOQ];
    double* J = s->f[ORC_CELL_JI];
    for (int n = 0; n < s->nc; ++n) {
        rho[n] /= dV;
        if (rho[n] < s->p.min_rho) { rho[n] = s->p.min_rho; cnt += 1; }
        for (int l = 0; l < 3; ++l) J[(size_t)n * 3 + l] /= dV;
    }
    s->counters[ORC_CNT_MIN_CELL_RHO] += cnt;
}

/* ParticleBoundaryCondHybrid::apply (src/particle_boundary_cond_hybrid.h:112-185):
 * empty exterior cells, remove particles inside the species' obstacle. */
void orc_apply_boundary_conditions(orc_sim* s, int sp) {
    const orc_species* S = &s->sp[sp];
    int64_t* cs = s->cstart[sp];
    double* st = s->pst[sp];
    int64_t out = 0;
    int64_t start0 = 0;
    for (int n = 0; n < s->nc; ++n) {
        int64_t b = start0, e = cs[n + 1];
        start0 = e;
        cs[n] = out;
        if (s->nflags[n] != ALL_NBRS) { /* exterior: delete all, counted as escapes */
            for (int64_t p = b; p < e; ++p) {
                const double* q = st + 7 * p;
                s->esc_w[sp] += q[6];
                s->esc_wv2[sp] += q[6] * (sqr(q[3]) + sqr(q[4]) + sqr(q[5]));
            }
            continue;
        }
        int ci, cj, ck;
        decompose(&s->g, n, &ci, &cj, &ck);
        const double xb = s->g.x0 + ci * s->g.dx, yb = s->g.y0 + cj * s->g.dx, zb = s->g.z0 + ck * s->g.dx;
        for (int64_t p = b; p < e; ++p) {
            const double* q = st + 7 * p;
            if (s->flag[ORC_INNER_PARTICLE][n]) {
                const double r2 = sqr(xb + q[0]) + sqr(yb + q[1]) + sqr(zb + q[2]);
                if (r2 < S->R2_obstacle) { /* impact */
                    s->imp_w[sp] += q[6];
                    s->imp_wv2[sp] += q[6] * (sqr(q[3]) + sqr(q[4]) + sqr(q[5]));
                    continue;
                }
            }
            if (out != p) memmove(st + 7 * out, q, 7 * sizeof(double));
            ++out;
        }
    }
    cs[s->nc] = out;
    s->np[sp] = out;
}

/* ------------------------------------------------------ field stencils */

/* face2Cell (src/hybrid_propagator.cpp:1043-1056). */
void orc_face2cell(orc_sim* s, const double* face, double* cell) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < s->nc; ++n) {
        if (s->nflags[n] != ALL_NBRS) continue;
        double aa[81];
        fill_nan(aa, 81);
        fetch(s, face, aa, n, 3);
        cell[n * 3 + 0] = 0.5 * (aa[AI(1, 1, 1) * 3 + 0] + aa[AI(0, 1, 1) * 3 + 0]);
        cell[n * 3 + 1] = 0.5 * (aa[AI(1, 1, 1) * 3 + 1] + aa[AI(1, 0, 1) * 3 + 1]);
        cell[n * 3 + 2] = 0.5 * (aa[AI(1, 1, 1) * 3 + 2] + aa[AI(1, 1, 0) * 3 + 2]);
    }
}

static inline int upper_missing(uint32_t nf) {
    return !(nf & X_POS) || !(nf & Y_POS) || !(nf & Z_POS);
}

/* cell2Node (src/hybrid_propagator.cpp:1059-1089). */
void orc_cell2node(orc_sim* s, const double* cell, double* node, int dim) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < s->nc; ++n) {
        if (upper_missing(s->nflags[n])) continue;
        double aa[81];
        fill_nan(aa, 81);
        fetch(s, cell, aa, n, dim);
        for (int l = 0; l < dim; ++l)
            node[(size_t)n * dim + l] = 0.125 * (aa[AI(1, 1, 1) * dim + l] + aa[AI(1, 1, 2) * dim + l] +
                                                 aa[AI(1, 2, 1) * dim + l] + aa[AI(2, 1, 1) * dim + l] +
                                                 aa[AI(1, 2, 2) * dim + l] + aa[AI(2, 2, 1) * dim + l] +
                                                 aa[AI(2, 1, 2) * dim + l] + aa[AI(2, 2, 2) * dim + l]);
    }
}

/* node2Cell (src/hybrid_propagator.cpp:1092-1112). */
void orc_node2cell(orc_sim* s, const double* node, double* cell) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < s->nc; ++n) {
        if (s->nflags[n] != ALL_NBRS) continue;
        double aa[81];
        fill_nan(aa, 81);
        fetch(s, node, aa, n, 3);
        for (int l = 0; l < 3; ++l)
            cell[n * 3 + l] = 0.125 * (aa[AI(1, 1, 1) * 3 + l] + aa[AI(0, 1, 1) * 3 + l] + aa[AI(0, 0, 1) * 3 + l] +
                                       aa[AI(1, 0, 1) * 3 + l] + aa[AI(1, 1, 0) * 3 + l] + aa[AI(1, 0, 0) * 3 + l] +
                                       aa[AI(0, 1, 0) * 3 + l] + aa[AI(0, 0, 0) * 3 + l]);
    }
}

/* neumannCell (src/hybrid_propagator.cpp:443-655): zero-gradient copies into
 * exterior cells; walls, then edges, then corners, later rules overwrite. */
void orc_neumann_cell(orc_sim* s, double* D, int dim) {
    static const int wall[6][4] = {
        /* mask-dir, src offset */
        {0, +1, 0, 0},  /* -x wall */
        {3, 0, -1, 0},  /* +y wall */
        {2, 0, +1, 0},  /* -y wall */
        {5, 0, 0, -1},  /* +z wall */
        {4, 0, 0, +1},  /* -z wall */
        {1, -1, 0, 0},  /* +x wall */
    };
    /* missing-direction pairs for the edges, in the reference's order */
    static const int edge[12][2] = {{0, 2}, {0, 3}, {1, 2}, {1, 3}, {0, 4}, {0, 5},
                                    {1, 4}, {1, 5}, {2, 4}, {2, 5}, {3, 4}, {3, 5}};
    static const int corner[8][3] = {{0, 2, 4}, {0, 2, 5}, {0, 3, 4}, {0, 3, 5},
                                     {1, 2, 4}, {1, 2, 5}, {1, 3, 4}, {1, 3, 5}};
    /* direction index: 0 -x, 1 +x, 2 -y, 3 +y, 4 -z, 5 +z; source offset is
     * opposite to the missing direction */
    static const int dsrc[6][3] = {{+1, 0, 0}, {-1, 0, 0}, {0, +1, 0}, {0, -1, 0}, {0, 0, +1}, {0, 0, -1}};
    const uint32_t dmask[6] = {X_NEG, X_POS, Y_NEG, Y_POS, Z_NEG, Z_POS};
    double aa[81];
    for (int e = 0; e < s->nexterior; ++e) {
        const int n = s->exterior[s->ext_reversed ? s->nexterior - 1 - e : e];
        fill_nan(aa, 81);
        fetch(s, D, aa, n, dim);
        const uint32_t nf = s->nflags[n];
        for (int w = 0; w < 6; ++w) {
            if (nf & dmask[wall[w][0]]) continue;
            const int m = AI(1 + wall[w][1], 1 + wall[w][2], 1 + wall[w][3]);
            for (int l = 0; l < dim; ++l) D[(size_t)n * dim + l] = aa[m * dim + l];
        }
        for (int ed = 0; ed < 12; ++ed) {
            const int d0 = edge[ed][0], d1 = edge[ed][1];
            if ((nf & dmask[d0]) || (nf & dmask[d1])) continue;
            const int m = AI(1 + dsrc[d0][0] + dsrc[d1][0], 1 + dsrc[d0][1] + dsrc[d1][1], 1 + dsrc[d0][2] + dsrc[d1][2]);
            for (int l = 0; l < dim; ++l) D[(size_t)n * dim + l] = aa[m * dim + l];
        }
        for (int c = 0; c < 8; ++c) {
            const int d0 = corner[c][0], d1 = corner[c][1], d2 = corner[c][2];
            if ((nf & dmask[d0]) || (nf & dmask[d1]) || (nf & dmask[d2])) continue;
            const int m = AI(1 + dsrc[d0][0] + dsrc[d1][0] + dsrc[d2][0], 1 + dsrc[d0][1] + dsrc[d1][1] + dsrc[d2][1],
                             1 + dsrc[d0][2] + dsrc[d1][2] + dsrc[d2][2]);
            for (int l = 0; l < dim; ++l) D[(size_t)n * dim + l] = aa[m * dim + l];
        }
    }
}

/* neumannFace (src/hybrid_propagator.cpp:659-908): tangential components
 * only on walls; per-component rules on edges and corners.  Each rule is
 * (missing mask, component, source buffer index) in the reference's order. */
typedef struct { uint32_t need; int comp; int a, b, c; } face_rule;
void orc_neumann_face(orc_sim* s, double* D) {
    const face_rule R[] = {
        /* walls */
        {X_NEG, 1, 2, 1, 1}, {X_NEG, 2, 2, 1, 1},
        {Y_POS, 0, 1, 0, 1}, {Y_POS, 2, 1, 0, 1},
        {Y_NEG, 0, 1, 2, 1}, {Y_NEG, 2, 1, 2, 1},
        {Z_POS, 0, 1, 1, 0}, {Z_POS, 1, 1, 1, 0},
        {Z_NEG, 0, 1, 1, 2}, {Z_NEG, 1, 1, 1, 2},
        {X_POS, 1, 0, 1, 1}, {X_POS, 2, 0, 1, 1},
        /* edges */
        {X_NEG | Y_NEG, 0, 1, 2, 1}, {X_NEG | Y_NEG, 1, 2, 1, 1}, {X_NEG | Y_NEG, 2, 2, 2, 1},
        {X_NEG | Y_POS, 0, 1, 0, 1}, {X_NEG | Y_POS, 2, 2, 0, 1},
        {X_POS | Y_NEG, 1, 0, 1, 1}, {X_POS | Y_NEG, 2, 0, 2, 1},
        {X_POS | Y_POS, 2, 0, 0, 1},
        {X_NEG | Z_NEG, 0, 1, 1, 2}, {X_NEG | Z_NEG, 1, 2, 1, 2}, {X_NEG | Z_NEG, 2, 2, 1, 1},
        {X_NEG | Z_POS, 0, 1, 1, 0}, {X_NEG | Z_POS, 1, 2, 1, 0},
        {X_POS | Z_NEG, 1, 0, 1, 2}, {X_POS | Z_NEG, 2, 0, 1, 1},
        {X_POS | Z_POS, 1, 0, 1, 0},
        {Y_NEG | Z_NEG, 0, 1, 2, 2}, {Y_NEG | Z_NEG, 1, 1, 1, 2}, {Y_NEG | Z_NEG, 2, 1, 2, 1},
        {Y_NEG | Z_POS, 0, 1, 2, 0}, {Y_NEG | Z_POS, 1, 1, 1, 0},
        {Y_POS | Z_NEG, 0, 1, 0, 2}, {Y_POS | Z_NEG, 2, 1, 0, 1},
        {Y_POS | Z_POS, 0, 1, 0, 0},
        /* corners */
        {X_NEG | Y_NEG | Z_NEG, 0, 1, 2, 2}, {X_NEG | Y_NEG | Z_NEG, 1, 2, 1, 2}, {X_NEG | Y_NEG | Z_NEG, 2, 2, 2, 1},
        {X_NEG | Y_NEG | Z_POS, 0, 1, 2, 0}, {X_NEG | Y_NEG | Z_POS, 1, 2, 1, 0},
        {X_NEG | Y_POS | Z_NEG, 0, 1, 0, 2}, {X_NEG | Y_POS | Z_NEG, 2, 2, 0, 1},
        {X_NEG | Y_POS | Z_POS, 0, 1, 0, 0},
        {X_POS | Y_NEG | Z_NEG, 1, 0, 1, 2}, {X_POS | Y_NEG | Z_NEG, 2, 0, 2, 1},
        {X_POS | Y_NEG | Z_POS, 1, 0, 1, 0},
        {X_POS | Y_POS | Z_NEG, 2, 0, 0, 1},
    };
    const int nr = (int)(sizeof(R) / sizeof(R[0]));
    double aa[81];
    for (int e = 0; e < s->nexterior; ++e) {
        const int n = s->exterior[s->ext_reversed ? s->nexterior - 1 - e : e];
        fill_nan(aa, 81);
        fetch(s, D, aa, n, 3);
        const uint32_t nf = s->nflags[n];
        for (int r = 0; r < nr; ++r) {
            if ((nf & R[r].need) != 0) continue; /* all named neighbours must be missing */
            D[(size_t)n * 3 + R[r].comp] = aa[AI(R[r].a, R[r].b, R[r].c) * 3 + R[r].comp];
        }
    }
}

/* setIMFCell (src/hybrid_propagator.cpp:910-977). */
void orc_set_imf_cell(orc_sim* s, double* cellB) {
    const uint32_t dm[6] = {X_POS, X_NEG, Y_POS, Y_NEG, Z_POS, Z_NEG};
    for (int e = 0; e < s->nexterior; ++e) {
        const int n = s->exterior[e];
        for (int w = 0; w < 6; ++w) {
            if ((s->nflags[n] & dm[w]) == 0 && s->p.imf_cell[w])
                for (int l = 0; l < 3; ++l) cellB[(size_t)n * 3 + l] = s->p.imf[l];
        }
    }
}

/* setIMFFace (src/hybrid_propagator.cpp:979-1040): tangential components. */
void orc_set_imf_face(orc_sim* s, double* faceB) {
    const uint32_t dm[6] = {X_POS, X_NEG, Y_POS, Y_NEG, Z_POS, Z_NEG};
    for (int e = 0; e < s->nexterior; ++e) {
        const int n = s->exterior[e];
        for (int w = 0; w < 6; ++w) {
            if ((s->nflags[n] & dm[w]) != 0 || !s->p.imf_face[w]) continue;
            const int axis = w / 2;
            for (int l = 0; l < 3; ++l)
                if (l != axis) faceB[(size_t)n * 3 + l] = s->p.imf[l];
        }
    }
}

/* calcNodeJ (src/hybrid_propagator.cpp:1449-1522): edge currents from the
 * curl of faceB, averaged to nodes, with the whistler-speed limiter. */
void orc_calc_node_J(orc_sim* s) {
    const double* faceB = s->f[ORC_FACE_B];
    const double* nodeB = s->f[ORC_NODE_B];
    const double* nodeRho = s->f[ORC_NODE_RHOQ];
    double* nodeJ = s->f[ORC_NODE_J];
    const double dx = s->g.dx;
    double cvw = 0;
#pragma omp parallel for schedule(static) reduction(+ : cvw)
    for (int n = 0; n < s->nc; ++n) {
        if (upper_missing(s->nflags[n])) continue;
        double aa[81];
        fill_nan(aa, 81);
        fetch(s, faceB, aa, n, 3);
#define F(a, b, c, l) aa[AI(1 + (a), 1 + (b), 1 + (c)) * 3 + (l)]
        const double ex1 = -F(0, 0, 0, 2) + F(0, 0, 0, 1) + F(0, 1, 0, 2) - F(0, 0, 1, 1);
        const double ex2 = -F(1, 0, 0, 2) + F(1, 0, 0, 1) + F(1, 1, 0, 2) - F(1, 0, 1, 1);
        const double ey1 = +F(0, 0, 0, 2) + F(0, 0, 1, 0) - F(1, 0, 0, 2) - F(0, 0, 0, 0);
        const double ey2 = +F(0, 1, 0, 2) + F(0, 1, 1, 0) - F(1, 1, 0, 2) - F(0, 1, 0, 0);
        const double ez1 = -F(0, 0, 0, 1) + F(0, 0, 0, 0) + F(1, 0, 0, 1) - F(0, 1, 0, 0);
        const double ez2 = -F(0, 0, 1, 1) + F(0, 0, 1, 0) + F(1, 0, 1, 1) - F(0, 1, 1, 0);
#undef F
        double d = 1.0, vw = 0.0;
        const double ne = nodeRho[n] / ORC_E;
        const double Btot = sqrt(sqr(nodeB[n * 3 + 0]) + sqr(nodeB[n * 3 + 1]) + sqr(nodeB[n * 3 + 2]));
        if (ne > 0.0 && dx > 0.0) vw = 2.0 * Btot * M_PI / (ORC_MU0 * ne * ORC_E * dx);
        if (vw > s->p.max_whistler && s->p.max_whistler > 0.0) { d = vw / s->p.max_whistler; cvw += 1; }
        const double a = 0.5 / (dx * ORC_MU0 * d);
        nodeJ[n * 3 + 0] = (ex1 + ex2) * a;
        nodeJ[n * 3 + 1] = (ey1 + ey2) * a;
        nodeJ[n * 3 + 2] = (ez1 + ez2) * a;
    }
    s->counters[ORC_CNT_MAX_VW] += cvw;
}

/* calcCellUe (src/hybrid_propagator.cpp:1342-1387). */
void orc_calc_cell_Ue(orc_sim* s) {
    const double* J = s->f[ORC_CELL_J];
    const double* Ji = s->f[ORC_CELL_JI];
    const double* rho = s->f[ORC_CELL_RHOQ];
    double* Ue = s->f[ORC_CELL_UE];
    double cnt = 0;
#pragma omp parallel for schedule(static) reduction(+ : cnt)
    for (int n = 0; n < s->nc; ++n) {
        if (s->nflags[n] != ALL_NBRS) continue;
        if (s->flag[ORC_INNER_FIELD][n]) {
            for (int l = 0; l < 3; ++l) Ue[n * 3 + l] = s->p.obstacle_ue[l];
            continue;
        }
        for (int l = 0; l < 3; ++l) {
            if (fabs(rho[n]) > 0) {
                if (s->p.use_hall) Ue[n * 3 + l] = -(J[n * 3 + l] - Ji[n * 3 + l]) / rho[n];
                else Ue[n * 3 + l] = Ji[n * 3 + l] / rho[n];
            } else {
                Ue[n * 3 + l] = 0.0;
            }
        }
        const double Ue2 = sqr(Ue[n * 3 + 0]) + sqr(Ue[n * 3 + 1]) + sqr(Ue[n * 3 + 2]);
        if (Ue2 > s->p.max_electron_speed2) {
            const double norm = sqrt(s->p.max_electron_speed2 / Ue2);
            for (int l = 0; l < 3; ++l) Ue[n * 3 + l] *= norm;
            cnt += 1;
        }
    }
    s->counters[ORC_CNT_MAX_CELL_UE] += cnt;
}

/* calcNodeUe (src/hybrid_propagator.cpp:1524-1584); clamps nodeRhoQi in place. */
void orc_calc_node_Ue(orc_sim* s) {
    double* rho = s->f[ORC_NODE_RHOQ];
    const double* Ji = s->f[ORC_NODE_JI];
    const double* J = s->f[ORC_NODE_J];
    double* Ue = s->f[ORC_NODE_UE];
    double crho = 0, cue = 0;
#pragma omp parallel for schedule(static) reduction(+ : crho, cue)
    for (int n = 0; n < s->nc; ++n) {
        if (upper_missing(s->nflags[n])) continue;
        if (s->flag[ORC_INNER_NODE][n]) {
            for (int l = 0; l < 3; ++l) Ue[n * 3 + l] = s->p.obstacle_ue[l];
            continue;
        }
        if (rho[n] < s->p.min_rho) { rho[n] = s->p.min_rho; crho += 1; }
        for (int l = 0; l < 3; ++l) {
            if (fabs(rho[n]) > 0) {
                if (s->p.use_hall) Ue[n * 3 + l] = -(J[n * 3 + l] - Ji[n * 3 + l]) / rho[n];
                else Ue[n * 3 + l] = Ji[n * 3 + l] / rho[n];
            } else {
                Ue[n * 3 + l] = 0.0;
            }
        }
        const double Ue2 = sqr(Ue[n * 3 + 0]) + sqr(Ue[n * 3 + 1]) + sqr(Ue[n * 3 + 2]);
        if (Ue2 > s->p.max_electron_speed2) {
            const double norm = sqrt(s->p.max_electron_speed2 / Ue2);
            for (int l = 0; l < 3; ++l) Ue[n * 3 + l] *= norm;
            cue += 1;
        }
    }
    s->counters[ORC_CNT_MIN_NODE_RHO] += crho;
    s->counters[ORC_CNT_MAX_NODE_UE] += cue;
}

/* upwindNodeB (src/hybrid_propagator.cpp:1268-1339). */
void orc_upwind_node_B(orc_sim* s) {
    const double* cellB = s->f[ORC_CELL_B];
    const double* Ue = s->f[ORC_NODE_UE];
    double* nodeB = s->f[ORC_NODE_B];
#pragma omp parallel for schedule(static)
    for (int n = 0; n < s->nc; ++n) {
        if (upper_missing(s->nflags[n])) continue;
        double aa[81];
        fill_nan(aa, 81);
        fetch(s, cellB, aa, n, 3);
        double xu = -0.5, yu = 0, zu = 0;
        const double ue = sqrt(sqr(Ue[n * 3 + 0]) + sqr(Ue[n * 3 + 1]) + sqr(Ue[n * 3 + 2]));
        if (ue > 0) {
            const double a = -0.5 / ue;
            xu = Ue[n * 3 + 0] * a; yu = Ue[n * 3 + 1] * a; zu = Ue[n * 3 + 2] * a;
        }
        /* the 8 cells around the node, in the reference's order:
         * 111,112,121,211,122,221,212,222 */
        static const int cc[8][3] = {{1, 1, 1}, {1, 1, 2}, {1, 2, 1}, {2, 1, 1}, {1, 2, 2}, {2, 2, 1}, {2, 1, 2}, {2, 2, 2}};
        double w[8], wsum = 0;
        for (int q = 0; q < 8; ++q) {
            const double xc = cc[q][0] == 1 ? -0.5 : +0.5, yc = cc[q][1] == 1 ? -0.5 : +0.5, zc = cc[q][2] == 1 ? -0.5 : +0.5;
            w[q] = 1 / sqrt(sqr(xc - xu) + sqr(yc - yu) + sqr(zc - zu));
        }
        wsum = w[0] + w[1] + w[2] + w[3] + w[4] + w[5] + w[6] + w[7];
        if (wsum > 0) {
            for (int l = 0; l < 3; ++l) {
                double acc = 0;
                for (int q = 0; q < 8; ++q) acc += w[q] * aa[AI(cc[q][0], cc[q][1], cc[q][2]) * 3 + l];
                nodeB[n * 3 + l] = acc / wsum;
            }
        } else {
            nodeB[n * 3 + 0] = nodeB[n * 3 + 1] = nodeB[n * 3 + 2] = 0.0;
        }
    }
}

/* calcNodeE (src/hybrid_propagator.cpp:1391-1446): E = -Ue x B + eta J. */
void orc_calc_node_E(orc_sim* s) {
    const double* Ue = s->f[ORC_NODE_UE];
    const double* B = s->f[ORC_NODE_B];
    const double* eta = s->f[ORC_NODE_ETA];
    const double* J = s->f[ORC_NODE_J];
    double* E = s->f[ORC_NODE_E];
    double cnt = 0;
#pragma omp parallel for schedule(static) reduction(+ : cnt)
    for (int n = 0; n < s->nc; ++n) {
        if (upper_missing(s->nflags[n])) continue;
        const int n3 = n * 3;
        E[n3 + 0] = -(Ue[n3 + 1] * B[n3 + 2] - Ue[n3 + 2] * B[n3 + 1]);
        E[n3 + 1] = -(Ue[n3 + 2] * B[n3 + 0] - Ue[n3 + 0] * B[n3 + 2]);
        E[n3 + 2] = -(Ue[n3 + 0] * B[n3 + 1] - Ue[n3 + 1] * B[n3 + 0]);
        E[n3 + 0] += eta[n] * J[n3 + 0];
        E[n3 + 1] += eta[n] * J[n3 + 1];
        E[n3 + 2] += eta[n] * J[n3 + 2];
        if (s->p.max_E2 > 0) {
            const double E2 = sqr(E[n3 + 0]) + sqr(E[n3 + 1]) + sqr(E[n3 + 2]);
            if (E2 > s->p.max_E2) {
                const double sc = sqrt(s->p.max_E2 / E2);
                E[n3 + 0] *= sc; E[n3 + 1] *= sc; E[n3 + 2] *= sc;
                cnt += 1;
            }
        }
    }
    s->counters[ORC_CNT_MAX_E] += cnt;
}

/* nodeAvg (src/hybrid_propagator.cpp:1115-1265).  The reference allocates its
 * gather buffer as `new Real[s*vectorDim] {initVal}`, which sets element 0 to
 * the sentinel and value-initialises the rest to zero; so only a missing
 * (-1,-1,-1) neighbour's x component is recognised as missing and the other
 * missing entries count as zeros with their full weight.  Restated as is. */
void orc_node_avg(orc_sim* s, const double* oldE, double* E) {
    const double initVal = DBL_MAX;
    const double* C = s->p.gauss_coeffs;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < s->nc; ++n) {
        if (upper_missing(s->nflags[n])) continue;
        double aa[81];
        for (int i = 0; i < 81; ++i) aa[i] = 0.0;
        aa[0] = initVal;
        fetch(s, oldE, aa, n, 3);
        for (int l = 0; l < 3; ++l) {
            double val[27], w[27];
            int zeroFound = 0;
            for (int c = 0; c < 3; ++c)
                for (int b = 0; b < 3; ++b)
                    for (int a = 0; a < 3; ++a) {
                        const int t = AI(a, b, c);
                        const int dist = (a != 1) + (b != 1) + (c != 1);
                        w[t] = C[dist];
                        val[t] = aa[t * 3 + l];
                        if (val[t] == initVal) { val[t] = 0.0; w[t] = 0.0; zeroFound = 1; }
                    }
            /* the reference's summation order (nodeEl111, 110, 011, 121, ...) */
            static const int ord[27][3] = {{1, 1, 1}, {1, 1, 0}, {0, 1, 1}, {1, 2, 1}, {2, 1, 1}, {1, 0, 1}, {1, 1, 2},
                                           {0, 1, 0}, {1, 0, 0}, {1, 2, 0}, {2, 1, 0}, {0, 0, 1}, {0, 2, 1}, {2, 2, 1},
                                           {2, 0, 1}, {0, 1, 2}, {1, 2, 2}, {2, 1, 2}, {1, 0, 2}, {0, 0, 2}, {0, 2, 2},
                                           {2, 0, 2}, {2, 2, 2}, {0, 0, 0}, {0, 2, 0}, {2, 0, 0}, {2, 2, 0}};
            if (zeroFound) {
                double Csum = 0;
                for (int q = 0; q < 27; ++q) Csum += w[AI(ord[q][0], ord[q][1], ord[q][2])];
                if (Csum > 0.0) for (int t = 0; t < 27; ++t) w[t] /= Csum;
            }
            double acc = 0;
            for (int q = 0; q < 27; ++q) {
                const int t = AI(ord[q][0], ord[q][1], ord[q][2]);
                acc += w[t] * val[t];
            }
            E[n * 3 + l] = acc;
        }
    }
}

/* faceCurl, Faraday mode (src/hybrid_propagator.cpp:1587-1671). */
void orc_face_curl(orc_sim* s, const double* node, double* face, double dt) {
    const double dx = s->g.dx;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < s->nc; ++n) {
        const uint32_t nf = s->nflags[n];
        int doX = 1, doY = 1, doZ = 1;
        if (nf != ALL_NBRS) {
            if (upper_missing(nf)) continue;
            if (!(nf & X_NEG)) doY = doZ = 0;
            if (!(nf & Y_NEG)) doX = doZ = 0;
            if (!(nf & Z_NEG)) doX = doY = 0;
        }
        double aa[81];
        fill_nan(aa, 81);
        fetch(s, node, aa, n, 3);
        const double n1x = aa[AI(0, 0, 1) * 3 + 0], n1y = aa[AI(0, 0, 1) * 3 + 1];
        const double n3x = aa[AI(0, 1, 0) * 3 + 0], n3z = aa[AI(0, 1, 0) * 3 + 2];
        const double n4x = aa[AI(0, 1, 1) * 3 + 0], n4y = aa[AI(0, 1, 1) * 3 + 1], n4z = aa[AI(0, 1, 1) * 3 + 2];
        const double n5x = aa[AI(1, 0, 1) * 3 + 0], n5y = aa[AI(1, 0, 1) * 3 + 1], n5z = aa[AI(1, 0, 1) * 3 + 2];
        const double n6y = aa[AI(1, 0, 0) * 3 + 1], n6z = aa[AI(1, 0, 0) * 3 + 2];
        const double n7x = aa[AI(1, 1, 0) * 3 + 0], n7y = aa[AI(1, 1, 0) * 3 + 1], n7z = aa[AI(1, 1, 0) * 3 + 2];
        const double n8x = aa[AI(1, 1, 1) * 3 + 0], n8y = aa[AI(1, 1, 1) * 3 + 1], n8z = aa[AI(1, 1, 1) * 3 + 2];
        if (doX) face[n * 3 + 0] += dt * (0.5 * (+n5z + n6z - n6y - n7y - n7z - n8z + n8y + n5y) / dx);
        if (doY) face[n * 3 + 1] += dt * (0.5 * (-n4x - n8x + n8z + n7z + n7x + n3x - n3z - n4z) / dx);
        if (doZ) face[n * 3 + 2] += dt * (0.5 * (-n1x - n5x - n5y - n8y + n8x + n4x + n4y + n1y) / dx);
    }
}

/* calcCellEp (src/hybrid_propagator.cpp:1674-1737). */
void orc_calc_cell_Ep(orc_sim* s) {
    const double* nodeRho = s->f[ORC_NODE_RHOQ];
    const double* cellRho = s->f[ORC_CELL_RHOQ];
    double* Ep = s->f[ORC_CELL_EP];
    const double dx = s->g.dx;
    const double coeff = s->p.e_pressure_coeff;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < s->nc; ++n) {
        if (s->nflags[n] != ALL_NBRS) continue;
        double aa[27];
        fill_nan(aa, 27);
        fetch(s, nodeRho, aa, n, 1);
        Ep[n * 3 + 0] = Ep[n * 3 + 1] = Ep[n * 3 + 2] = 0.0;
        if (s->flag[ORC_INNER_CELLEP][n]) continue;
        const double xP = 0.25 * (aa[AI(1, 1, 1)] + aa[AI(1, 1, 0)] + aa[AI(1, 0, 0)] + aa[AI(1, 0, 1)]);
        const double xN = 0.25 * (aa[AI(0, 1, 1)] + aa[AI(0, 1, 0)] + aa[AI(0, 0, 0)] + aa[AI(0, 0, 1)]);
        const double yP = 0.25 * (aa[AI(1, 1, 1)] + aa[AI(1, 1, 0)] + aa[AI(0, 1, 0)] + aa[AI(0, 1, 1)]);
        const double yN = 0.25 * (aa[AI(1, 0, 1)] + aa[AI(1, 0, 0)] + aa[AI(0, 0, 0)] + aa[AI(0, 0, 1)]);
        const double zP = 0.25 * (aa[AI(1, 1, 1)] + aa[AI(1, 0, 1)] + aa[AI(0, 0, 1)] + aa[AI(0, 1, 1)]);
        const double zN = 0.25 * (aa[AI(1, 1, 0)] + aa[AI(1, 0, 0)] + aa[AI(0, 0, 0)] + aa[AI(0, 1, 0)]);
        const double gx = (xP - xN) / dx, gy = (yP - yN) / dx, gz = (zP - zN) / dx;
        if (!s->p.isothermal) {
            Ep[n * 3 + 0] = -coeff * gx; Ep[n * 3 + 1] = -coeff * gy; Ep[n * 3 + 2] = -coeff * gz;
        } else if (fabs(cellRho[n]) > 0) {
            Ep[n * 3 + 0] = -coeff * gx / cellRho[n];
            Ep[n * 3 + 1] = -coeff * gy / cellRho[n];
            Ep[n * 3 + 2] = -coeff * gz / cellRho[n];
        }
    }
}

/* propagateB (src/hybrid_propagator.cpp:42-440), variant USE_NODE_UE +
 * USE_EDGE_J (src/Makefile:1-7); temporal averages are left out. */
void orc_propagate_B(orc_sim* s) {
    double** f = s->f;
    orc_face2cell(s, f[ORC_FACE_B], f[ORC_CELL_B]);
    orc_neumann_cell(s, f[ORC_CELL_B], 3);
    orc_neumann_cell(s, f[ORC_CELL_JI], 3);
    orc_neumann_cell(s, f[ORC_CELL_RHOQ], 1);
    orc_set_imf_cell(s, f[ORC_CELL_B]);
    orc_cell2node(s, f[ORC_CELL_B], f[ORC_NODE_B], 3);
    orc_cell2node(s, f[ORC_CELL_RHOQ], f[ORC_NODE_RHOQ], 1);
    orc_cell2node(s, f[ORC_CELL_JI], f[ORC_NODE_JI], 3);
    orc_neumann_face(s, f[ORC_FACE_B]);
    orc_set_imf_face(s, f[ORC_FACE_B]);
    orc_calc_node_J(s);
    orc_node2cell(s, f[ORC_NODE_J], f[ORC_CELL_J]);
    orc_calc_cell_Ue(s);
    orc_neumann_cell(s, f[ORC_CELL_UE], 3);
    orc_calc_node_Ue(s);
    orc_upwind_node_B(s);
    orc_calc_node_E(s);
    for (int i = 0; i < s->p.n_linear_smoothings; ++i) {
        orc_node2cell(s, f[ORC_NODE_E], f[ORC_CELL_J]);
        orc_neumann_cell(s, f[ORC_CELL_J], 3);
        orc_cell2node(s, f[ORC_CELL_J], f[ORC_NODE_E], 3);
        memset(f[ORC_CELL_J], 0, sizeof(double) * 3 * (size_t)s->nc);
    }
    if (s->p.use_gaussian) {
        double* old = (double*)malloc(sizeof(double) * 3 * (size_t)s->nc);
        memcpy(old, f[ORC_NODE_E], sizeof(double) * 3 * (size_t)s->nc);
        orc_node_avg(s, old, f[ORC_NODE_E]);
        free(old);
    }
    orc_face_curl(s, f[ORC_NODE_E], f[ORC_FACE_B], s->p.dt);
    if (s->p.use_e_pressure) orc_calc_cell_Ep(s);
}

/* ------------------------------------------------------------ injection */

/* Philox4x32-10: ten rounds of two 32x32->64 multiplications with the key
 * bumped by the Weyl constants between rounds (Salmon et al., SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 53 random bits -> [0,1) */
static inline double u53(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

typedef struct {
    uint32_t key[2], ctr[4];
    double su, sg;
    int hu, hg;
} orc_stream;

static void stream_init(orc_stream* st, const uint32_t key[2], uint32_t c0, uint32_t c1, uint32_t c2) {
    st->key[0] = key[0]; st->key[1] = key[1];
    st->ctr[0] = c0; st->ctr[1] = c1; st->ctr[2] = c2; st->ctr[3] = 0;
    st->hu = st->hg = 0;
}

static double stream_uniform(orc_stream* st) {
    if (st->hu) { st->hu = 0; return st->su; }
    uint32_t o[4];
    orc_philox4x32_10(st->ctr, st->key, o);
    st->ctr[3]++;
    st->su = u53(o[2], o[3]);
    st->hu = 1;
    return u53(o[0], o[1]);
}

/* sampleGaussian (src/particle_injector.cpp:59-80): polar method, the second
 * deviate kept for the next call (per stream here).  r2 == 0 is also
 * rejected (log(0)). */
static double stream_gauss(orc_stream* st) {
    if (st->hg) { st->hg = 0; return st->sg; }
    double x, y, r2;
    do {
        x = 2 * stream_uniform(st) - 1;
        y = 2 * stream_uniform(st) - 1;
        r2 = x * x + y * y;
    } while (r2 >= 1.0 || r2 == 0.0);
    const double fac = sqrt(-2.0 * log(r2) / r2);
    st->sg = y * fac;
    st->hg = 1;
    return x * fac;
}

/* constants of sampleDerivGaussian (src/particle_injector.cpp:97-106) */
static void deriv_gauss_consts(double x0, double* xm, double* c, double* d, double* cxm) {
    *xm = 0.5 * (x0 + sqrt(x0 * x0 + 4.0));
    *c = 1.0 / (exp(-0.5 * x0 * x0) + x0 * sqrt(0.5 * M_PI) * erfc(-x0 * (1.0 / sqrt(2.0))));
    *d = sqr(x0 - *xm);
    *cxm = *c * *xm;
}

/* sampleDerivGaussian (src/particle_injector.cpp:97-114): f(x) ~ max(0,x)
 * exp(-(x-x0)^2/2) by rejection from the unit Gaussian at its mode xm. */
static double stream_deriv_gauss(orc_stream* st, double x0, double xm, double c, double d, double cxm) {
    for (;;) {
        const double x = xm + stream_gauss(st);
        if (x < 0) continue;
        const double majorant = cxm * exp(-0.5 * (sqr(x - xm) + d));
        const double pdf = c * x * exp(-0.5 * sqr(x - x0));
        if (stream_uniform(st) * majorant > pdf) continue;
        return x;
    }
}

/* probRound (src/particle_injector.cpp:41-47) */
static int prob_round(double x, double r) {
    if (x <= 0) return 0;
    const int f = (int)floor(x);
    return (r < x - f) ? f + 1 : f;
}

void orc_rng_sample(uint64_t seed, int which, double x0, int64_t n, double* out) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_stream st;
    stream_init(&st, key, 0, 0, 0);
    double xm = 0, c = 0, d = 0, cxm = 0;
    if (which == 2) deriv_gauss_consts(x0, &xm, &c, &d, &cxm);
    for (int64_t q = 0; q < n; ++q) {
        switch (which) {
        case 0: out[q] = stream_uniform(&st); break;
        case 1: out[q] = stream_gauss(&st); break;
        case 2: out[q] = stream_deriv_gauss(&st, x0, xm, c, d, cxm); break;
        default: out[q] = prob_round(x0, stream_uniform(&st)); break;
        }
    }
}

int orc_set_injector(orc_sim* s, int sp, const orc_injector* in) {
    if (sp < 0 || sp >= s->nspecies) return -1;
    struct orc_inj* I = &s->inj[sp];
    free(I->cell_rate);
    memset(I, 0, sizeof(*I));
    if (!in || in->kind == ORC_INJ_NONE) return 0;
    I->key[0] = (uint32_t)in->seed;
    I->key[1] = (uint32_t)(in->seed >> 32);
    if (!(in->temperature > 0)) return -2;
    I->vth = sqrt(ORC_KB * in->temperature / s->sp[sp].m);
    if (in->kind == ORC_INJ_SOLAR_WIND) {
        /* InjectorSolarWind::initialize (src/particle_injector.cpp:588-682) */
        int nz = 0, axis = -1;
        for (int l = 0; l < 3; ++l)
            if (in->velocity[l] != 0.0) { ++nz; axis = l; }
        if (nz != 1 || !(in->ppc > 0) || !(in->density > 0)) return -2;
        const double U = fabs(in->velocity[axis]);
        I->axis = axis;
        I->pos_wall = in->velocity[axis] < 0;
        I->rate = in->ppc * U * s->p.dt / s->g.dx;
        I->x0 = U / I->vth;
        deriv_gauss_consts(I->x0, &I->xm, &I->c, &I->d, &I->cxm);
    } else if (in->kind == ORC_INJ_UNIFORM) {
        /* InjectorUniform::initialize (src/particle_injector.cpp:267-345) */
        if (!(in->ppc > 0) || !(in->density > 0)) return -2;
        for (int l = 0; l < 3; ++l) {
            I->U[l] = in->velocity[l];
            if (!(in->box[2 * l] < in->box[2 * l + 1])) return -2;
        }
        memcpy(I->box, in->box, sizeof(I->box));
        I->rate = in->ppc;
    } else if (in->kind == ORC_INJ_EXOSPHERE) {
        if (!in->cell_rate) return -2;
        I->cell_rate = (double*)malloc(sizeof(double) * s->nc);
        memcpy(I->cell_rate, in->cell_rate, sizeof(double) * s->nc);
        I->R2_exobase = in->exobase_radius > 0 ? sqr(in->exobase_radius) : 0.0;
    } else {
        return -2;
    }
    I->kind = in->kind;
    return 0;
}

/* checkIfInjectionCell{X,Y,Z}{Pos,Neg} (src/particle_injector.cpp:392-454):
 * the wall neighbour is missing, the four transverse neighbours exist. */
static int sw_injection_cell(const orc_sim* s, const struct orc_inj* I, int n) {
    static const uint32_t pos[3] = {X_POS, Y_POS, Z_POS}, neg[3] = {X_NEG, Y_NEG, Z_NEG};
    const uint32_t nf = s->nflags[n];
    if (nf & (I->pos_wall ? pos[I->axis] : neg[I->axis])) return 0;
    for (int l = 0; l < 3; ++l) {
        if (l == I->axis) continue;
        if (!(nf & pos[l]) || !(nf & neg[l])) return 0;
    }
    return 1;
}

/* The cell centre lies in the uniform injector's box (src/particle_injector.cpp:207-210). */
static int uniform_cell(const orc_sim* s, const struct orc_inj* I, int n) {
    int i, j, k;
    decompose(&s->g, n, &i, &j, &k);
    const double c[3] = {s->g.x0 + (i + 0.5) * s->g.dx, s->g.y0 + (j + 0.5) * s->g.dx, s->g.z0 + (k + 0.5) * s->g.dx};
    for (int l = 0; l < 3; ++l)
        if (c[l] < I->box[2 * l] || c[l] > I->box[2 * l + 1]) return 0;
    return 1;
}

/* One injection cell's new particles (cell-local metres, state[7]) written
 * to out; returns their number.  Stream of the cell's count: particle index
 * 0xFFFFFFFF; stream of trial particle m: index m. */
static int inject_cell(const orc_sim* s, int sp, int n, uint32_t step, double* out, int max_out) {
    const struct orc_inj* I = &s->inj[sp];
    const double dx = s->g.dx;
    orc_stream st;
    stream_init(&st, I->key, (uint32_t)n, step, 0xFFFFFFFFu);
    const double rate = I->kind == ORC_INJ_EXOSPHERE ? I->cell_rate[n] : I->rate;
    const int N = prob_round(rate, stream_uniform(&st));
    int q = 0;
    for (int m = 0; m < N; ++m) {
        stream_init(&st, I->key, (uint32_t)n, step, (uint32_t)m);
        double r[3], v[3];
        if (I->kind == ORC_INJ_SOLAR_WIND) {
            /* initParticleCrdVel{X,Y,Z}{Pos,Neg} (src/particle_injector.cpp:464-535):
             * on the wall face, uniform across it; flux-weighted normal speed */
            const int a = I->axis, t1 = (a + 1) % 3, t2 = (a + 2) % 3;
            const int l1 = t1 < t2 ? t1 : t2, l2 = t1 < t2 ? t2 : t1;
            r[a] = I->pos_wall ? 0.0 : dx;
            r[l1] = stream_uniform(&st) * dx;
            r[l2] = stream_uniform(&st) * dx;
            const double vn = I->vth * stream_deriv_gauss(&st, I->x0, I->xm, I->c, I->d, I->cxm);
            v[a] = I->pos_wall ? -vn : vn;
            v[l1] = I->vth * stream_gauss(&st);
            v[l2] = I->vth * stream_gauss(&st);
        } else if (I->kind == ORC_INJ_UNIFORM) {
            /* InjectorUniform::injectParticles (src/particle_injector.cpp:187-252):
             * uniform in the cell shrunk by epsCell, drifting Maxwellian */
            const double eps = 1.0e-4;
            for (int l = 0; l < 3; ++l) r[l] = 0.5 * dx + (1.0 - eps) * dx * (stream_uniform(&st) - 0.5);
            for (int l = 0; l < 3; ++l) v[l] = I->U[l] + I->vth * stream_gauss(&st);
        } else {
            /* InjectorExosphere::injectParticles (src/particle_injector.cpp:1522-1591):
             * uniform in the cell shrunk by epsCell, discarded below the exobase */
            int i, j, k;
            decompose(&s->g, n, &i, &j, &k);
            const double eps = 1.0e-4;
            for (int l = 0; l < 3; ++l) r[l] = 0.5 * dx + (1.0 - eps) * dx * (stream_uniform(&st) - 0.5);
            const double gx = (s->g.x0 + i * dx) + r[0], gy = (s->g.y0 + j * dx) + r[1],
                         gz = (s->g.z0 + k * dx) + r[2];
            if (gx * gx + gy * gy + gz * gz < I->R2_exobase) continue;
            for (int l = 0; l < 3; ++l) v[l] = I->vth * stream_gauss(&st);
        }
        if (q < max_out) {
            double* o = out + 7 * q;
            o[0] = r[0]; o[1] = r[1]; o[2] = r[2];
            o[3] = v[0]; o[4] = v[1]; o[5] = v[2];
            o[6] = s->sp[sp].w;
        }
        ++q;
    }
    return q;
}

/* ParticleInjectorBase::inject: every listed cell, in cell order; the new
 * particles go to the end of their cell's list. */
void orc_inject(orc_sim* s, int sp) {
    if (sp < 0 || sp >= s->nspecies) return;
    struct orc_inj* I = &s->inj[sp];
    if (I->kind == ORC_INJ_NONE || s->timestep <= 0) return;
    if (I->kind == ORC_INJ_UNIFORM && s->timestep != 1) return; /* :181 */
    const uint32_t step = (uint32_t)s->timestep;
    const int nc = s->nc;
    int* cnt = (int*)calloc((size_t)nc, sizeof(int));
    int64_t total = 0, cap = 0;
    for (int n = 0; n < nc; ++n) {
        if (I->kind == ORC_INJ_SOLAR_WIND && !sw_injection_cell(s, I, n)) continue;
        if (I->kind == ORC_INJ_EXOSPHERE && !(I->cell_rate[n] > 0)) continue;
        if (I->kind == ORC_INJ_UNIFORM && !uniform_cell(s, I, n)) continue;
        cnt[n] = inject_cell(s, sp, n, step, NULL, 0);
        total += cnt[n];
        if (cnt[n] > cap) cap = cnt[n];
    }
    if (total == 0) { free(cnt); return; }
    const int64_t np = s->np[sp];
    double* nst = (double*)malloc(sizeof(double) * 7 * (size_t)(np + total));
    double* tmp = (double*)malloc(sizeof(double) * 7 * (size_t)cap);
    int64_t* cs = s->cstart[sp];
    int64_t out = 0;
    for (int n = 0; n < nc; ++n) {
        const int64_t b = cs[n], e = cs[n + 1];
        cs[n] = out;
        if (e > b) memcpy(nst + 7 * out, s->pst[sp] + 7 * b, sizeof(double) * 7 * (size_t)(e - b));
        out += e - b;
        if (cnt[n]) {
            inject_cell(s, sp, n, step, tmp, cnt[n]);
            memcpy(nst + 7 * out, tmp, sizeof(double) * 7 * (size_t)cnt[n]);
            out += cnt[n];
            for (int q = 0; q < cnt[n]; ++q) { /* src/particle_injector.cpp:573-575 */
                const double* o = tmp + 7 * q;
                s->inj_w[sp] += o[6];
                s->inj_macro[sp] += 1.0;
                s->inj_wv2[sp] += o[6] * (sqr(o[3]) + sqr(o[4]) + sqr(o[5]));
            }
        }
    }
    cs[nc] = out;
    free(s->pst[sp]);
    s->pst[sp] = nst;
    s->np[sp] = out;
    I->injected += total;
    free(tmp);
    free(cnt);
}

void orc_set_timestep(orc_sim* s, int64_t t) { s->timestep = t; }
int64_t orc_timestep(const orc_sim* s) { return s->timestep; }
int64_t orc_injected(const orc_sim* s, int sp) { return (sp >= 0 && sp < s->nspecies) ? s->inj[sp].injected : -1; }

/* propagate() (src/user.cpp:118-134) without logging and detectors: the
 * timestep advances, then push all, clear, migrate, accumulate all,
 * finalise, particle boundary conditions, injection, field update. */
void orc_step(orc_sim* s) {
    s->timestep++;
    for (int sp = 0; sp < s->nspecies; ++sp) orc_push(s, sp);
    orc_clear_accumulation(s);
    for (int sp = 0; sp < s->nspecies; ++sp) orc_migrate(s, sp);
    for (int sp = 0; sp < s->nspecies; ++sp) orc_accumulate(s, sp);
    orc_add_remote_updates(s);
    for (int sp = 0; sp < s->nspecies; ++sp) orc_apply_boundary_conditions(s, sp);
    for (int sp = 0; sp < s->nspecies; ++sp) orc_inject(s, sp);
    orc_propagate_B(s);
}

/* ------------------------------------------------------------ logs */

/* Constants of the logs (CODATA; Corsair's table is not in the reference tree). */
#define ORC_MP   1.672621716e-27
#define ORC_ME   9.10938215e-31
#define ORC_EPS0 8.854187817e-12

void orc_log_begin(orc_sim* s, double t) {
    s->log_t0 = t;
    memcpy(s->log_cnt0, s->counters, sizeof(s->counters));
    for (int sp = 0; sp < s->nspecies; ++sp) {
        s->esc_w[sp] = s->esc_wv2[sp] = s->imp_w[sp] = s->imp_wv2[sp] = 0.0;
        s->inj_w[sp] = s->inj_wv2[sp] = s->inj_macro[sp] = s->maxvi[sp] = 0.0;
    }
}

/* logWriteParticleField (src/diagnostics.h:407-813) with logCalcParticle
 * (:187-251) and logCalcField (:254-404): the values of one line of each
 * pop log (17 per species) and of field.log (41), then the counters restart
 * (:778-799).  Column 31 repeats the node-Ue counter, as :589 does. */
void orc_log_values(orc_sim* s, double t, int include_inner, double field[41], double* pop) {
    const int nc = s->nc;
    const double dx = s->g.dx, dV = dx * dx * dx, dt = s->p.dt;
    const double Dt = t - s->log_t0;
    double* rhoM = (double*)calloc((size_t)nc, sizeof(double));
    for (int sp = 0; sp < s->nspecies; ++sp) {
        double Nr = 0, sx = 0, sy = 0, sz = 0, sv = 0, sv2 = 0, mx = 0;
        for (int n = 0; n < nc; ++n)
            for (int64_t p = s->cstart[sp][n]; p < s->cstart[sp][n + 1]; ++p) {
                const double* q = s->pst[sp] + 7 * p;
                const double w = q[6];
                Nr += w;
                sx += w * q[3]; sy += w * q[4]; sz += w * q[5];
                const double v2 = sqr(q[3]) + sqr(q[4]) + sqr(q[5]), v = sqrt(v2);
                sv += w * v; sv2 += w * v2;
                if (v > mx) mx = v;
                rhoM[n] += w * s->sp[sp].m;
            }
        const double m = s->sp[sp].m;
        double* o = pop + 17 * sp;
        o[0] = t; o[1] = Nr; o[2] = (double)s->np[sp];
        if (Nr > 0) { o[3] = sx / Nr; o[4] = sy / Nr; o[5] = sz / Nr; o[6] = sv / Nr; }
        else { o[3] = o[4] = o[5] = o[6] = 0.0; }
        o[7] = 0.5 * m * sv2; o[8] = mx;
        if (Dt > 0) {
            o[9] = s->esc_w[sp] / Dt; o[10] = s->imp_w[sp] / Dt; o[11] = s->inj_w[sp] / Dt;
            o[12] = s->inj_macro[sp] / Dt * dt;
            o[13] = 0.5 * m * s->esc_wv2[sp] / Dt; o[14] = 0.5 * m * s->imp_wv2[sp] / Dt;
            o[15] = 0.5 * m * s->inj_wv2[sp] / Dt; o[16] = s->maxvi[sp] / Dt * dt;
        } else {
            for (int q = 9; q < 17; ++q) o[q] = 0.0;
        }
    }
    const double minRhoM = s->p.min_rho * ORC_MP / ORC_E;
    for (int n = 0; n < nc; ++n) {
        rhoM[n] /= dV;
        if (rhoM[n] < minRhoM) rhoM[n] = minRhoM;
    }
    double N = 0, sB[3] = {0, 0, 0}, sBt = 0, mB = 0, sDiv = 0, mDiv = 0, mDivPerB = 0, sB2 = 0;
    double sJ[3] = {0, 0, 0}, sJt = 0, mJ = 0, sEp[3] = {0, 0, 0}, sEpt = 0, mEp = 0, sEp2 = 0;
    double sE[3] = {0, 0, 0}, sEt = 0, mE = 0, sE2 = 0, sCE2 = 0;
    double minDe = DBL_MAX, maxDe = -DBL_MAX, minDi = DBL_MAX, maxDi = -DBL_MAX, minTL = DBL_MAX, maxVA = 0, maxUe = 0;
    const double* fB = s->f[ORC_FACE_B];
    const double* Ji = s->f[ORC_CELL_JI];
    const double* Ep = s->f[ORC_CELL_EP];
    const double* nE = s->f[ORC_NODE_E];
    const double* Ue = s->f[ORC_CELL_UE];
    const double* rq = s->f[ORC_CELL_RHOQ];
    for (int n = 0; n < nc; ++n) {
        if (s->nflags[n] != ALL_NBRS) continue;
        if (!include_inner && s->flag[ORC_INNER_FIELD][n]) continue;
        const int nxm = nbr(&s->g, n, -1, 0, 0), nym = nbr(&s->g, n, 0, -1, 0), nzm = nbr(&s->g, n, 0, 0, -1);
        double divB = ((fB[3 * n + 0] - fB[3 * nxm + 0]) + (fB[3 * n + 1] - fB[3 * nym + 1]) +
                       (fB[3 * n + 2] - fB[3 * nzm + 2])) / dx;
        divB = fabs(divB);
        const double Bx = 0.5 * (fB[3 * n + 0] + fB[3 * nxm + 0]);
        const double By = 0.5 * (fB[3 * n + 1] + fB[3 * nym + 1]);
        const double Bz = 0.5 * (fB[3 * n + 2] + fB[3 * nzm + 2]);
        const double B2 = sqr(Bx) + sqr(By) + sqr(Bz), Bt = sqrt(B2);
        const double dpb = Bt > 0.0 ? divB / Bt : 0.0;
        sB[0] += Bx; sB[1] += By; sB[2] += Bz; sBt += Bt; sDiv += divB; sB2 += B2;
        if (Bt > mB) mB = Bt;
        if (fabs(divB) > fabs(mDiv)) mDiv = divB;
        if (fabs(dpb) > fabs(mDivPerB)) mDivPerB = dpb;
        const double Jt = sqrt(sqr(Ji[3 * n]) + sqr(Ji[3 * n + 1]) + sqr(Ji[3 * n + 2]));
        for (int l = 0; l < 3; ++l) sJ[l] += Ji[3 * n + l];
        sJt += Jt;
        if (Jt > mJ) mJ = Jt;
        const double Ep2 = sqr(Ep[3 * n]) + sqr(Ep[3 * n + 1]) + sqr(Ep[3 * n + 2]), Ept = sqrt(Ep2);
        for (int l = 0; l < 3; ++l) sEp[l] += Ep[3 * n + l];
        sEpt += Ept; sEp2 += Ep2;
        if (Ept > mEp) mEp = Ept;
        const double E2 = sqr(nE[3 * n]) + sqr(nE[3 * n + 1]) + sqr(nE[3 * n + 2]), Et = sqrt(E2);
        for (int l = 0; l < 3; ++l) sE[l] += nE[3 * n + l];
        sEt += Et; sE2 += E2;
        if (Et > mE) mE = Et;
        const double ux = Ue[3 * n], uy = Ue[3 * n + 1], uz = Ue[3 * n + 2];
        const double cEx = -(uy * Bz - uz * By), cEy = -(uz * Bx - ux * Bz), cEz = -(ux * By - uy * Bx);
        sCE2 += sqr(cEx) + sqr(cEy) + sqr(cEz);
        const double Uet = sqrt(sqr(ux) + sqr(uy) + sqr(uz));
        if (rq[n] > 0.0) {
            const double de = sqrt(ORC_ME / (ORC_MU0 * ORC_E * rq[n]));
            if (de > maxDe) maxDe = de;
            if (de < minDe) minDe = de;
        }
        if (Bt > 0.0) {
            const double tL = 2.0 * M_PI * ORC_MP / (ORC_E * Bt);
            if (tL < minTL) minTL = tL;
        }
        if (rhoM[n] > 0.0) {
            const double di = sqrt(sqr(ORC_MP) / (ORC_MU0 * sqr(ORC_E) * rhoM[n]));
            const double vA = Bt / sqrt(ORC_MU0 * rhoM[n]);
            if (di > maxDi) maxDi = di;
            if (di < minDi) minDi = di;
            if (vA > maxVA) maxVA = vA;
        }
        if (Uet > maxUe) maxUe = Uet;
        N += 1.0;
    }
    free(rhoM);
    field[0] = t;
    for (int l = 0; l < 3; ++l) field[1 + l] = N > 0 ? sB[l] / N : 0.0;
    field[4] = N > 0 ? sBt / N : 0.0;
    field[5] = mB;
    field[6] = N > 0 ? sDiv / N : 0.0;
    field[7] = mDiv; field[8] = dx * mDivPerB; field[9] = sB2 * dV / (2.0 * ORC_MU0);
    for (int l = 0; l < 3; ++l) field[10 + l] = N > 0 ? sJ[l] / N : 0.0;
    field[13] = N > 0 ? sJt / N : 0.0;
    field[14] = mJ;
    for (int l = 0; l < 3; ++l) field[15 + l] = N > 0 ? sEp[l] / N : 0.0;
    field[18] = N > 0 ? sEpt / N : 0.0;
    field[19] = mEp; field[20] = sEp2 * dV * ORC_EPS0 / 2.0;
    for (int l = 0; l < 3; ++l) field[21 + l] = N > 0 ? sE[l] / N : 0.0;
    field[24] = N > 0 ? sEt / N : 0.0;
    field[25] = mE; field[26] = sE2 * dV * ORC_EPS0 / 2.0; field[27] = sCE2 * dV * ORC_EPS0 / 2.0;
    const double* c = s->counters;
    const double* c0 = s->log_cnt0;
    const double cnt[6] = {c[ORC_CNT_MAX_CELL_UE] - c0[ORC_CNT_MAX_CELL_UE], c[ORC_CNT_MAX_NODE_UE] - c0[ORC_CNT_MAX_NODE_UE],
                           c[ORC_CNT_MAX_NODE_UE] - c0[ORC_CNT_MAX_NODE_UE], c[ORC_CNT_MAX_E] - c0[ORC_CNT_MAX_E],
                           c[ORC_CNT_MIN_CELL_RHO] - c0[ORC_CNT_MIN_CELL_RHO],
                           c[ORC_CNT_MIN_NODE_RHO] - c0[ORC_CNT_MIN_NODE_RHO]};
    for (int q = 0; q < 6; ++q) field[28 + q] = Dt > 0 ? cnt[q] / Dt * dt : 0.0;
    field[34] = minDe; field[35] = maxDe; field[36] = minDi; field[37] = maxDi;
    field[38] = minTL; field[39] = maxVA; field[40] = maxUe;
    orc_log_begin(s, t);
}

/* logCalcParticle sums (src/diagnostics.h:209-220). */
void orc_diagnostics(const orc_sim* s, int sp, double out[5]) {
    double a = 0, bx = 0, by = 0, bz = 0, e = 0;
    for (int64_t p = 0; p < s->np[sp]; ++p) {
        const double* q = s->pst[sp] + 7 * p;
        a += q[6];
        bx += q[6] * q[3]; by += q[6] * q[4]; bz += q[6] * q[5];
        e += q[6] * (sqr(q[3]) + sqr(q[4]) + sqr(q[5]));
    }
    out[0] = a; out[1] = bx; out[2] = by; out[3] = bz; out[4] = e;
}
