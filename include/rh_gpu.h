/*
 * rh_gpu.h -- C ABI of the MI355X-native RHybrid hot path.
 *
 * One context (rh_ctx) owns the device state of one simulation domain on one
 * GPU: the grid arrays, the particle stores of every species and the scratch
 * of the sort and deposit.  All calls enqueue on the context's HIP stream;
 * upload/download/diagnostics calls synchronise.  Every function returns
 * RH_OK (0) or a negative rh_status; rh_last_error() gives the message.  The
 * library never aborts.
 *
 * The operators replace the reference's Corsair operator surface one for one
 * (file:line in the reference tree are given at each declaration).  Arrays
 * cross the boundary in the reference's layouts: FP64, cell n = i + nx*(j +
 * ny*k), vector components interleaved n*3+l (src/user.cpp:839-927), and
 * particles as (cell, AoS state[7] = {x,y,z,vx,vy,vz,w}) with cell-local
 * coordinates in metres (src/particle_definition.h:29-38).
 */
#ifndef RH_GPU_H
#define RH_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rh_ctx rh_ctx;

typedef enum rh_status {
    RH_OK = 0,
    RH_ERR_ARG = -1,       /* bad argument (null pointer, id or size out of range) */
    RH_ERR_HIP = -2,       /* a HIP runtime call failed */
    RH_ERR_STATE = -3,     /* call not valid in the current state */
    RH_ERR_CAPACITY = -4,  /* a particle store or index would overflow */
    RH_ERR_NO_DEVICE = -5  /* no usable GPU */
} rh_status;

typedef enum rh_precision { RH_FP32 = 0, RH_FP64 = 1 } rh_precision;

/* Grid of nx*ny*nz cells including the exterior layer of non-periodic
 * dimensions (the LogicallyCartesian grids of the reference's example
 * configs); cubic cells of edge dx; (x0,y0,z0) is the lower corner of cell
 * (0,0,0).  tile[] is the cell tile of the sort and deposit (0 = choose);
 * each must divide its dimension. */
typedef struct rh_grid_desc {
    int32_t nx, ny, nz;
    int32_t periodic[3];
    double dx;
    double x0, y0, z0;
    int32_t tile[3];
} rh_grid_desc;

/* The Hybrid:: parameters the path reads (src/hybrid.h:253-300), in SI, as
 * src/user.cpp:1565-1807 derives them. */
typedef struct rh_params {
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
    double imf[3];                /* Hybrid::IMFB{x,y,z} */
    int32_t imf_cell[6];          /* Hybrid::IMFBoundaryCellB, +x -x +y -y +z -z */
    int32_t imf_face[6];          /* Hybrid::IMFBoundaryFaceB */
    double obstacle_ue[3];        /* Hybrid::fieldObstacleUe */
    int32_t initial_flow_through; /* Hybrid::useInitialFlowThrough */
    int32_t use_gravity;          /* Hybrid::useGravity */
    double GMdt;                  /* Hybrid::GMdt */
} rh_params;

/* Species (src/particle_species.h:30-46): charge and mass in SI, the
 * macroparticle weight, squared obstacle radius (<0: none), the accumulate and
 * accelerate flags, and the particle capacity to reserve (0: 1.5x the first
 * upload). */
typedef struct rh_species_desc {
    double q, m, w;
    double R2_obstacle;
    int32_t accumulate, accelerate;
    int64_t capacity;
} rh_species_desc;

/* Field ids: the grid arrays of src/user.cpp:839-927. */
enum rh_field_id {
    RH_FACE_B = 0, RH_CELL_B, RH_CELL_J, RH_CELL_UE, RH_CELL_JI, RH_CELL_EP, RH_CELL_RHOQ,
    RH_NODE_RHOQ, RH_NODE_E, RH_NODE_B, RH_NODE_J, RH_NODE_UE, RH_NODE_JI, RH_NODE_ETA,
    RH_N_FIELDS
};
/* Inner-boundary flag ids (src/user.cpp:1882-1912). */
enum rh_flag_id {
    RH_INNER_FIELD = 0, RH_INNER_NODE, RH_INNER_PARTICLE, RH_INNER_CELLEP, RH_N_FLAGS
};
/* Constraint counters, as Hybrid::logCounter* (src/hybrid.h:306-330). */
enum rh_counter_id {
    RH_CNT_MAX_VI = 0, RH_CNT_MIN_CELL_RHO, RH_CNT_MAX_CELL_UE, RH_CNT_MIN_NODE_RHO,
    RH_CNT_MAX_NODE_UE, RH_CNT_MAX_E, RH_CNT_MAX_VW, RH_CNT_DROPPED, RH_CNT_DENOISE_FALLBACK, RH_N_COUNTERS
};

/* ---- context -------------------------------------------------------- */

/* Replaces userEarlyInitialization + the array allocation of
 * userLateInitialization (src/user.cpp:160-185, 839-1050): picks the device
 * (the current HIP device), allocates grid arrays (zeroed, as
 * src/user.cpp:1853-1870) and creates the stream. */
int rh_create(const rh_grid_desc* grid, const rh_params* params, int32_t precision, rh_ctx** out);
int rh_destroy(rh_ctx* ctx);
const char* rh_last_error(const rh_ctx* ctx);
/* Species::readParameters + particle list creation (src/particle_species.cpp:38-107). */
int rh_add_species(rh_ctx* ctx, const rh_species_desc* desc, int32_t* id);

/* ---- data in and out (ParGrid getUserData* accessors) ---------------- */

int rh_upload_field(rh_ctx* ctx, int32_t field_id, const double* data);
int rh_download_field(rh_ctx* ctx, int32_t field_id, double* data);
int rh_upload_flags(rh_ctx* ctx, int32_t flag_id, const uint8_t* data);
/* Replace the particles of a species; cells must be valid, coordinates are
 * cell-local metres in [0,dx), w is ignored (the species weight is used). */
int rh_upload_particles(rh_ctx* ctx, int32_t species, int64_t n, const int64_t* cell, const double* state7);
int rh_num_particles(rh_ctx* ctx, int32_t species, int64_t* n);
/* Live particles in the store's order: by tile, then by cell inside the tile,
 * then by half-cell octant; w is the species weight. */
int rh_download_particles(rh_ctx* ctx, int32_t species, int64_t capacity, int64_t* cell, double* state7);

/* ---- the operators ---------------------------------------------------- */

/* setupGetFields (src/hybrid_propagator.cpp:1861-1876): faceB halo exchange;
 * nothing to exchange inside one GPU. */
int rh_setup_get_fields(rh_ctx* ctx);
/* BorisBuneman::propagateCell for every cell of a species
 * (src/particle_propagator_boris_buneman.h:196-207, :64-194). */
int rh_propagate(rh_ctx* ctx, int32_t species);
/* ParticleListSkeleton::waitParticleSends (src/user.cpp:123, Corsair): the
 * cell migration, done as a deterministic counting sort by cell. */
int rh_wait_particle_sends(rh_ctx* ctx, int32_t species);
/* Accumulator::clearAccumulationArrays (src/particle_accumulator.cpp:337-362). */
int rh_clear_accumulation_arrays(rh_ctx* ctx);
/* Accumulator::accumulateBoundaryCells + accumulateInnerCells
 * (src/particle_accumulator.h:36-37, .cpp:202-254). */
int rh_accumulate(rh_ctx* ctx, int32_t species);
/* Accumulator::sendUpdates + wait + addRemoteUpdates
 * (src/particle_accumulator.cpp:381-409, 260-335): the fixed-order FP64 sum of
 * tile ghost contributions, /dV and the rho floor. */
int rh_add_remote_updates(rh_ctx* ctx);
/* ParticleBoundaryCondHybrid::apply (src/particle_boundary_cond_hybrid.h:55-187). */
int rh_apply_boundary_conditions(rh_ctx* ctx, int32_t species);
/* propagateB (src/hybrid_propagator.cpp:42-440). */
int rh_propagate_B(rh_ctx* ctx);
/* propagate() (src/user.cpp:118-134) without logging, nsteps times: each
 * step advances the timestep by one, then push, migration, deposit, particle
 * boundary conditions, injection and field update.  Every grid array holds
 * the reference's values when the call returns; the field updates of the
 * steps before the last write only the arrays a step reads (faceB, cellUe,
 * cellEp; cellB too with a denoiser attached), the last writes all.
 * Synchronises at the end; returns RH_ERR_CAPACITY when a particle store
 * overflowed in these steps. */
int rh_step(rh_ctx* ctx, int32_t nsteps);
/* No counterpart in the reference (its ranks push one species after the
 * other): with enable != 0, rh_step runs each species' push, migration and
 * deposit on a stream of its own, side by side, and joins them before the
 * ghost sums.  The results are bit for bit those of the default one-stream
 * order.  For species whose load is concentrated in a few crowded tiles (a
 * photoion cloud), whose launches leave most of the device idle.  Off by
 * default; the per-kernel times of rh_kernel_times then include the overlap. */
int rh_set_species_streams(rh_ctx* ctx, int32_t enable);

/* ---- injection (src/particle_injector.cpp) ----------------------------- */

/* Injector kinds: the solar-wind inflow through one outer wall
 * (InjectorSolarWind, src/particle_injector.cpp:380-700) and photoions from a
 * per-cell rate table (InjectorExosphere, :1489-1840). */
typedef enum rh_injector_kind {
    RH_INJ_NONE = 0, RH_INJ_SOLAR_WIND = 1, RH_INJ_EXOSPHERE = 2, RH_INJ_UNIFORM = 3
} rh_injector_kind;

/* Random numbers: Philox4x32-10 with key = seed and counter = (cell,
 * timestep, particle, block), so a step's new particles do not depend on
 * the order of evaluation (Corsair's sequential generator, which the
 * reference uses, is not reproduced).  Temperatures give vth = sqrt(kB T/m)
 * with kB = 1.3806505e-23. */
typedef struct rh_injector_desc {
    int32_t kind;              /* rh_injector_kind */
    uint64_t seed;
    double temperature;        /* [K] */
    /* RH_INJ_SOLAR_WIND: inflow through the wall the bulk velocity points
     * away from (velocity (U,0,0), U < 0: the +x wall), into the wall's
     * non-edge exterior cells, probRound(ppc*|U|*dt/dx) particles per cell per
     * step on the wall face with a flux-weighted normal speed
     * (src/particle_injector.cpp:392-577,681).  The species weight must be
     * density*dx^3/ppc (:682). */
    double velocity[3];        /* [m/s], exactly one non-zero component */
    double density;            /* [m^-3] */
    double ppc;                /* macroparticles per cell of the undisturbed wind */
    /* RH_INJ_EXOSPHERE: probRound(cell_rate[n]) trial particles per cell per
     * step, uniform in the cell shrunk by epsCell = 1e-4, discarded below the
     * exobase, Maxwellian velocities (:1522-1591).  cell_rate (nc values,
     * copied) is cellExosphere after its normalisation (:1772-1781). */
    const double* cell_rate;   /* macroparticles per cell per dt */
    double exobase_radius;     /* [m] */
    /* RH_INJ_UNIFORM (InjectorUniform, :159-345): at timestep 1 only,
     * probRound(ppc) particles in every cell whose centre lies in box (xmin
     * xmax ymin ymax zmin zmax [m]), uniform in the cell shrunk by epsCell,
     * velocity `velocity` + a Maxwellian of `temperature`; the species weight
     * must be density*dx^3/ppc. */
    double box[6];
} rh_injector_desc;

/* Attach (or with NULL / RH_INJ_NONE detach) the injector of a species
 * (ParticleInjectorBase::initialize). */
int rh_set_injector(rh_ctx* ctx, int32_t species, const rh_injector_desc* desc);
/* ParticleInjectorBase::inject (src/particle_injector.cpp:540-551) for the
 * current timestep; nothing at timestep <= 0.  The new particles are pushed
 * and deposited from the next step on, as in the reference's step order. */
int rh_inject_particles(rh_ctx* ctx, int32_t species);
/* Simulation::timestep, the injectors' counter (0 after rh_create). */
int rh_set_timestep(rh_ctx* ctx, int64_t timestep);
int rh_get_timestep(rh_ctx* ctx, int64_t* timestep);
/* macroparticles injected so far (Hybrid::logCounterParticleInjectMacroparticles) */
int rh_injected(rh_ctx* ctx, int32_t species, int64_t* n);

/* ---- learned moment denoiser (include/rh_denoise.h) -------------------- */

/* No counterpart in the reference: replaces the deposited ion moments
 * (cellRhoQi, cellJi) by the denoiser's output, with cellB as the magnetic
 * input channels, for callers that run at low ppc.  Call between
 * rh_add_remote_updates and rh_propagate_B.  Fallback bound: a cell whose
 * denoised charge density differs from the binomially filtered deposit
 * (weights 1/4, 1/2, 1/4 per axis) by more than max_rel_change * |filtered|
 * keeps its deposited moments and is counted in RH_CNT_DENOISE_FALLBACK; the
 * rho floor is applied again afterwards.
 * Needs nx % 16 == 0, ny % 4 == 0, nz % 4 == 0 and a denoiser created for at
 * least nx*ny*nz voxels. */
typedef struct rh_denoiser rh_denoiser;
int rh_denoise_moments(rh_ctx* ctx, rh_denoiser* dn, double max_rel_change);
/* Attach a denoiser that rh_step applies before every field update (with the
 * same fallback bound); NULL detaches.  The denoiser's work is then ordered
 * on the context's stream. */
int rh_set_denoiser(rh_ctx* ctx, rh_denoiser* dn, double max_rel_change);

/* ---- diagnostics ------------------------------------------------------ */

/* logCalcParticle sums in FP64 (src/diagnostics.h:209-220):
 * out = {sum w, sum w vx, sum w vy, sum w vz, sum w v^2}. */
int rh_diagnostics(rh_ctx* ctx, int32_t species, double out[5]);
int rh_counters(rh_ctx* ctx, double out[RH_N_COUNTERS]);
int rh_synchronize(rh_ctx* ctx);

/* ---- logs (src/diagnostics.h:187-816, src/user.cpp:2147-2256) ----------- */

/* The values of one line of the reference's log files at simulation time t
 * (logWriteParticleField): field[41] = the columns of field.log, pop[17 *
 * nspecies] = the columns of each pop*.log, all in FP64 (particle sums over
 * live particles including injected ones not yet pushed; field sums over
 * cells whose 26 neighbours exist, inner-boundary cells only with
 * include_inner_cells, as Hybrid::includeInnerCellsInFieldLog).  Rates divide
 * the counters gathered since the previous line (or rh_log_begin) by the time
 * since then; the counters then restart.  Column 31 repeats the node-Ue
 * counter, as the reference does (src/diagnostics.h:589). */
int rh_log_values(rh_ctx* ctx, double t, int32_t include_inner_cells, double field[41], double* pop);
/* Restart the log counters at time t (Hybrid::logCounterTimeStart); a new
 * context starts them at t = 0. */
int rh_log_begin(rh_ctx* ctx, double t);
/* Create field.log and pop<NNN>_<name>.log in `directory` with the
 * reference's headers (src/user.cpp:2147-2231), `precision` digits in
 * scientific notation with signs (Hybrid.log_precision, default 10). */
int rh_log_open(rh_ctx* ctx, const char* directory, const char* const* species_names, int32_t precision,
                int32_t include_inner_cells, double max_B_stop);
/* Append one line to every log file at time t (rh_log_values); returns
 * RH_ERR_STATE after writing when max(|faceB|) exceeds max_B_stop > 0, where
 * the reference stops the run (Hybrid::maxBStopRun). */
int rh_log_write(rh_ctx* ctx, double t);

/* ---- output (UserDataOP::writeData, src/operator_userdata.cpp:55-311) ---- */

/* Write a VLSV file of the selected cell variables at time t: the grid
 * arrays by name ("faceB", "cellRhoQi", "cellB", "cellJ", "cellUe",
 * "cellJi", "cellEp" (with electron pressure), "nodeRhoQi", "nodeE",
 * "nodeB", "nodeJ", "nodeUe", "nodeJi", "nodeEta"; FP64, interleaved n*3+l),
 * the inner-boundary flags (one byte per cell), "ne" (cellRhoQi/e),
 * "cellDivB", "cellNPles", and "n", "v", "T" for every output population
 * (n_<pop>, v_<pop>, T_<pop>; population_names[s] names species s's
 * population, species with one name are combined).  The file carries the
 * mesh "SpatialGrid" (MESH_BBOX, MESH_NODE_CRDS_X/Y/Z, CellID = cell index)
 * and the parameters time, timestep, dt and the box extent, as the
 * reference's reader (tools/rhybridpy.py) expects.  With temporal averages
 * on (rh_set_temporal_averages), "cellBAverage", "n_ave" and "v_ave" write
 * the averages since the previous file (cellBAverage, n_<pop>_ave,
 * v_<pop>_ave, src/operator_userdata.cpp:156-232) and restart them. */
int rh_write_vlsv(rh_ctx* ctx, const char* filename, double t, int64_t timestep, const char* const* variables,
                  int32_t nvariables, const char* const* population_names);
/* WRITE_GRID_TEMPORAL_AVERAGES (src/Makefile:6, on in the reference's
 * default build): with enable, every field update adds cellB to a running
 * FP64 sum (src/hybrid_propagator.cpp:157-167) and every accumulation adds
 * each species' charge and current (before /dV) to its own
 * (src/particle_accumulator.cpp:148-151, 287-300); the sums and their count
 * start at zero.  Off (the default here) frees them. */
int rh_set_temporal_averages(rh_ctx* ctx, int32_t enable);

/* ---- measurement -------------------------------------------------------- */

/* Per-kernel timing with HIP events on the context's stream (the profile::
 * regions of src/hybrid_propagator.cpp:33-38).  When enabled, every launch is
 * bracketed by events; rh_kernel_times returns, for the kernel name, the
 * accumulated milliseconds and the number of launches since the last reset.
 * A launch made for one species (push, sort, move, deposit, boundary
 * conditions, injection) also counts under "<name>:<species id>", e.g.
 * "move:1". */
int rh_set_timing(rh_ctx* ctx, int32_t enable);
/* Restrict the timing to the launches of one kernel name (NULL or "": every
 * launch).  Each timed launch adds two event records to the stream, which
 * costs a few microseconds of device time per launch; a measurement of the
 * whole step times only the kernel it needs.  An extension of the timing
 * above, with no counterpart in the reference. */
int rh_set_timing_only(rh_ctx* ctx, const char* kernel);
int rh_kernel_times(rh_ctx* ctx, const char* kernel, double* ms, int64_t* launches);
int rh_reset_timing(rh_ctx* ctx);
/* Sorted-store layout facts for roofline accounting: out = {n_tiles,
 * tile_x, tile_y, tile_z, n_cells, particle bytes per particle}. */
int rh_layout_info(rh_ctx* ctx, int64_t out[6]);

#ifdef __cplusplus
}
#endif
#endif
