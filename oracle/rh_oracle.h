/*
 * rh_oracle.h -- FP64 CPU restatement of RHybrid's per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path in
 * rhybrid_amd/ is compared against (tests/, __graft_entry__.smoke(), and the
 * cpu_baseline leg of bench.py).  Nothing in the product path links or loads it.
 *
 * It restates, in the reference's own data layout (AoS Particle<double>
 * {x,y,z,vx,vy,vz,w} stored per cell with cell-local coordinates, grid arrays
 * interleaved n*3+l), the following reference code:
 *   Boris-Buneman push       src/particle_propagator_boris_buneman.h:64-207
 *   field gather             src/hybrid_propagator.cpp:1740-1778,1838-1891
 *   CIC accumulation         src/particle_accumulator.cpp:43-156,260-362
 *   field propagator         src/hybrid_propagator.cpp:42-1737
 *   particle BCs             src/particle_boundary_cond_hybrid.h:112-185
 *   step order               src/user.cpp:118-134
 *   conservation sums        src/diagnostics.h:209-220
 *   solar-wind injection     src/particle_injector.cpp:41-114,380-700
 *   exosphere injection      src/particle_injector.cpp:1489-1840
 *   particle and field logs  src/diagnostics.h:187-816
 * Cell-to-cell particle migration lives in the un-vendored Corsair
 * ParticleListSkeleton; it is re-specified here (see orc_migrate).
 */
#ifndef RH_ORACLE_H
#define RH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid: nx*ny*nz cells (the exterior layer of a non-periodic dimension is
 * part of the count, as in the reference's LogicallyCartesian grids).  Cell
 * n = i + nx*(j + ny*k); (x0,y0,z0) is the lower corner of cell (0,0,0). */
typedef struct orc_grid {
    int32_t nx, ny, nz;
    int32_t periodic[3];
    double dx;
    double x0, y0, z0;
} orc_grid;

/* Parameters of the hot path (the Hybrid:: statics the path reads,
 * src/hybrid.h:253-300, already converted to SI as src/user.cpp:1565-1807
 * does). */
typedef struct orc_params {
    double dt;
    int32_t use_hall;             /* Hybrid::useHallElectricField */
    int32_t use_e_pressure;       /* Hybrid::useElectronPressure */
    int32_t isothermal;           /* Hybrid::useIsothermalElectrons */
    double e_pressure_coeff;      /* Hybrid::electronPressureCoeff */
    double max_ion_speed2;        /* Hybrid::maxIonSpeed2 */
    double max_electron_speed2;   /* Hybrid::maxElectronSpeed2 */
    double min_rho;               /* Hybrid::minIonChargeDensity [C/m^3] */
    double max_E2;                /* Hybrid::maxElectricField2 (0: off) */
    double max_whistler;          /* Hybrid::maxWhistlerSpeed (0: off) */
    int32_t n_linear_smoothings;  /* Hybrid::N_linearSmoothingsNodeE */
    int32_t use_gaussian;         /* Hybrid::useGaussianSmoothingNodeE */
    double gauss_coeffs[4];       /* Hybrid::gaussianSmoothingCoeffsNodeE */
    double imf[3];                /* Hybrid::IMFBx,IMFBy,IMFBz */
    int32_t imf_cell[6];          /* IMFBoundaryCellB: +x -x +y -y +z -z */
    int32_t imf_face[6];          /* IMFBoundaryFaceB */
    double obstacle_ue[3];        /* Hybrid::fieldObstacleUe */
    int32_t initial_flow_through; /* Hybrid::useInitialFlowThrough */
    int32_t use_gravity;          /* Hybrid::useGravity */
    double GMdt;                  /* Hybrid::GMdt */
} orc_params;

/* Species (src/particle_species.h:30-46).  w is the macroparticle weight the
 * injectors assign (n*dV/ppc, src/particle_injector.cpp:305,682). */
typedef struct orc_species {
    double q, m, w;
    double R2_obstacle;           /* <0: no obstacle */
    int32_t accumulate, accelerate;
} orc_species;

enum orc_field_id {
    ORC_FACE_B = 0, ORC_CELL_B, ORC_CELL_J, ORC_CELL_UE, ORC_CELL_JI, ORC_CELL_EP,
    ORC_CELL_RHOQ, ORC_NODE_RHOQ, ORC_NODE_E, ORC_NODE_B, ORC_NODE_J, ORC_NODE_UE,
    ORC_NODE_JI, ORC_NODE_ETA, ORC_N_FIELDS
};
enum orc_flag_id {
    ORC_INNER_FIELD = 0, ORC_INNER_NODE, ORC_INNER_PARTICLE, ORC_INNER_CELLEP, ORC_N_FLAGS
};
enum orc_counter_id {
    ORC_CNT_MAX_VI = 0, ORC_CNT_MIN_CELL_RHO, ORC_CNT_MAX_CELL_UE, ORC_CNT_MIN_NODE_RHO,
    ORC_CNT_MAX_NODE_UE, ORC_CNT_MAX_E, ORC_CNT_MAX_VW, ORC_N_COUNTERS
};

typedef struct orc_sim orc_sim;

orc_sim* orc_create(const orc_grid* g, const orc_params* p);
void     orc_destroy(orc_sim* s);
int      orc_add_species(orc_sim* s, const orc_species* sp);
int      orc_num_cells(const orc_sim* s);
int      orc_field_dim(int field_id);
double*  orc_field(orc_sim* s, int field_id);
uint8_t* orc_flag(orc_sim* s, int flag_id);
void     orc_set_threads(orc_sim* s, int nthreads);

/* Particles cross the boundary as (cell, AoS state[7]) with cell-local
 * coordinates in metres; they are stored sorted by cell (stable). */
int      orc_set_particles(orc_sim* s, int species, int64_t n, const int64_t* cell, const double* state7);
int64_t  orc_num_particles(const orc_sim* s, int species);
int      orc_get_particles(const orc_sim* s, int species, int64_t* cell, double* state7);
const int64_t* orc_cell_start(const orc_sim* s, int species);

/* The operators, named after the reference code they restate. */
void orc_push(orc_sim* s, int species);                 /* BorisBuneman::propagateCell */
void orc_migrate(orc_sim* s, int species);              /* Corsair particle migration */
void orc_clear_accumulation(orc_sim* s);                /* Accumulator::clearAccumulationArrays */
void orc_accumulate(orc_sim* s, int species);           /* Accumulator::accumulate{Boundary,Inner}Cells */
void orc_add_remote_updates(orc_sim* s);                /* Accumulator::addRemoteUpdates */
void orc_apply_boundary_conditions(orc_sim* s, int species); /* ParticleBoundaryCondHybrid::apply */
void orc_propagate_B(orc_sim* s);                       /* propagateB */
void orc_step(orc_sim* s);                              /* propagate(), src/user.cpp:118-134;
                                                           advances the timestep first */

/* Sub-steps of propagateB exposed for unit checks. */
void orc_face2cell(orc_sim* s, const double* face, double* cell);
void orc_cell2node(orc_sim* s, const double* cell, double* node, int dim);
void orc_node2cell(orc_sim* s, const double* node, double* cell);
void orc_neumann_cell(orc_sim* s, double* cell, int dim);
void orc_neumann_face(orc_sim* s, double* face);
void orc_face_curl(orc_sim* s, const double* node, double* face, double dt);
void orc_set_imf_cell(orc_sim* s, double* cellB);
void orc_set_imf_face(orc_sim* s, double* faceB);
/* the stencils below read and write the simulation's own field arrays */
void orc_calc_node_J(orc_sim* s);
void orc_calc_cell_Ue(orc_sim* s);
void orc_calc_node_Ue(orc_sim* s);
void orc_upwind_node_B(orc_sim* s);
void orc_calc_node_E(orc_sim* s);
void orc_node_avg(orc_sim* s, const double* oldE, double* E);
void orc_calc_cell_Ep(orc_sim* s);

/* ---- injection (src/particle_injector.cpp) ----------------------------
 *
 * The reference draws from Corsair's sequential generator, which is not in
 * the reference tree; the restatement draws from Philox4x32-10 (Salmon,
 * Moraes, Dror and Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11) with key = seed and counter = (cell, timestep, particle, block), so
 * that any order of evaluation gives the same particles.  Two doubles of 53
 * random bits per generator call. */
enum { ORC_INJ_NONE = 0, ORC_INJ_SOLAR_WIND = 1, ORC_INJ_EXOSPHERE = 2, ORC_INJ_UNIFORM = 3 };
typedef struct orc_injector {
    int32_t kind;
    uint64_t seed;
    double temperature;        /* vth = sqrt(kB T / m) */
    double velocity[3];        /* solar wind: bulk velocity, one non-zero component */
    double density;            /* solar wind: n [m^-3] */
    double ppc;                /* solar wind: macroparticles per cell */
    const double* cell_rate;   /* exosphere: nc values, macroparticles per cell per dt (copied) */
    double exobase_radius;     /* exosphere: trial positions below it are discarded */
    double box[6];             /* uniform: xmin xmax ymin ymax zmin zmax of the filled cell centres */
} orc_injector;

void    orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* n draws from one stream (counter (0,0,0,block)): which = 0 uniform [0,1),
 * 1 sampleGaussian, 2 sampleDerivGaussian(x0), 3 probRound(x0). */
void    orc_rng_sample(uint64_t seed, int which, double x0, int64_t n, double* out);
int     orc_set_injector(orc_sim* s, int species, const orc_injector* inj);
/* ParticleInjectorBase::inject for the current timestep (nothing at <= 0). */
void    orc_inject(orc_sim* s, int species);
void    orc_set_timestep(orc_sim* s, int64_t t);
int64_t orc_timestep(const orc_sim* s);
/* macroparticles injected so far per species */
int64_t orc_injected(const orc_sim* s, int species);

/* ---- logs (src/diagnostics.h:187-816, src/user.cpp:2147-2256) ----------- */
/* Start the log counters at time t (logCounterTimeStart). */
void    orc_log_begin(orc_sim* s, double t);
/* One line of every pop log (17 values per species, into pop) and of
 * field.log (41 values) at time t; the log counters restart at t. */
void    orc_log_values(orc_sim* s, double t, int include_inner, double field[41], double* pop);

/* Field gather for one point (cell-local metres) as the push sees it. */
void orc_get_fields(orc_sim* s, int64_t cell, const double r[3], double B[3], double Ue[3], double Ep[3]);

/* FP64 conservation sums of one species (src/diagnostics.h:209-220):
 * out = {sum w, sum w*vx, sum w*vy, sum w*vz, sum w*v^2}. */
void orc_diagnostics(const orc_sim* s, int species, double out[5]);
double orc_counter(const orc_sim* s, int counter_id);
void   orc_reset_counters(orc_sim* s);
/* Reverse the order in which exterior cells are visited by the Neumann
 * boundary loops (used by a test showing the result is order independent). */
void   orc_set_exterior_order_reversed(orc_sim* s, int reversed);

#ifdef __cplusplus
}
#endif
#endif
