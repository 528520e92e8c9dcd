/*
 * kestrel_step.h -- C-ABI of the MI355X (gfx950) finite-volume morphodynamic
 * time step.
 *
 * The library replaces the body of Kestrel's IntegrateTo
 * (src/TimeStepper.f90:116-277).  A Fortran host keeps calling Run
 * (src/TimeStepper.f90:73-113) once per output interval; Run calls
 * kst_integrate_to instead of IntegrateTo and afterwards downloads the state
 * with kst_get_state into grid%tileContainer, so OutputSolutionData,
 * CalculateVolume and OutputAggregateData (src/Output.f90) read the same data
 * as before.
 *
 * Conventions follow the reference's own C bridges (src/GeoTiffRead.f90:43-100,
 * src/utm.f90:55-100): plain pointers and sizes, the caller owns every host
 * array, an int status (0 = success, negative = fatal) that the Fortran side
 * turns into FatalErrorMessage -> exit(1) (src/Messages.f90:73-89).
 * One host thread per handle; every call blocks until host-visible results
 * are ready.
 *
 * Grid layout at the boundary (all arrays row-major, x fastest):
 *   cells     NX*NY            with NX = nx_tiles*nx_per_tile, NY likewise;
 *                              cell (I,J) of the global grid is cell
 *                              (ii,jj) = (I mod nx_per_tile, J mod ny_per_tile)
 *                              of tile TileID(I/nx_per_tile+1, J/ny_per_tile+1)
 *                              (src/Grid.f90:249-256), 0-based here.
 *   vertices  (NX+1)*(NY+1)    vertex (I,J) is the south-west corner of cell (I,J).
 *   tiles     nx_tiles*ny_tiles tile t = ti + tj*nx_tiles (TileID - 1).
 */
#ifndef KESTREL_STEP_H
#define KESTREL_STEP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enumerations: the runtime-selected closures (src/Closures.f90:67-151,
 *      src/Limiters.f90:68-75, bound from the input deck in
 *      src/Parameters.f90:175-420 and src/SolverSettings.f90:102-126) ---- */
/* Boundary conditions (RunSet%bcs, src/DomainSettings.f90:123-137):
 * halt      a tile on the domain edge may not become active: fatal
 *           (src/UpdateTiles.f90:61-65);
 * periodic  tiles on opposite domain edges are neighbours
 *           (SetPeriodicBoundaryConds, src/UpdateTiles.f90:521-566), edge
 *           tiles may be active, ghosts have no corner tiles (:405-409);
 * dirichlet edge tiles are never activated (:66-68) and ghost tiles on the
 *           domain edge hold the boundary data bc_hn, bc_u, bc_v, bc_psi
 *           (SetDomainBoundaryData, :590-643);
 * sponge    edge tiles are never activated, edge ghosts hold the default
 *           dry data, and ApplySpongeLayer (src/TimeStepper.f90:879-920)
 *           damps the explicit RHS of active tiles without a neighbour --
 *           which no active tile is, edge tiles never being active. */
enum { KST_BC_HALT = 0, KST_BC_PERIODIC = 1, KST_BC_DIRICHLET = 2, KST_BC_SPONGE = 3 };
enum { KST_LIMITER_MINMOD1 = 0, KST_LIMITER_MINMOD2 = 1, KST_LIMITER_VANALBADA = 2,
       KST_LIMITER_WENO = 3, KST_LIMITER_NONE = 4 };
enum { KST_DRAG_CHEZY = 0, KST_DRAG_COULOMB = 1, KST_DRAG_VOELLMY = 2, KST_DRAG_POULIQUEN = 3,
       KST_DRAG_EDWARDS2019 = 4, KST_DRAG_MANNING = 5, KST_DRAG_VARIABLE = 6 };
enum { KST_EROSION_NONE = 0, KST_EROSION_SIMPLE = 1, KST_EROSION_FLUID = 2,
       KST_EROSION_GRANULAR = 3, KST_EROSION_MIXED = 4 };
enum { KST_DEPOSITION_NONE = 0, KST_DEPOSITION_SIMPLE = 1, KST_DEPOSITION_SPEARMAN_MANNING = 2 };
enum { KST_TRANSITION_NONE = 0, KST_TRANSITION_SMOOTH = 1, KST_TRANSITION_STEP = 2 };
enum { KST_DAMPING_NONE = 0, KST_DAMPING_TANH = 1, KST_DAMPING_RAT3 = 2 };
enum { KST_SWITCH_TANH = 0, KST_SWITCH_RAT3 = 1, KST_SWITCH_COS = 2, KST_SWITCH_EQUAL = 3,
       KST_SWITCH_ZERO = 4, KST_SWITCH_ONE = 5, KST_SWITCH_LINEAR = 6, KST_SWITCH_STEP = 7 };
/* Storage/arithmetic precision of the device state. */
enum { KST_PREC_FP64 = 0, KST_PREC_MIXED = 1 };
/* Per-tile state (TileType%TileOn / %isGhostTile, src/Grid.f90:59-60). */
enum { KST_TILE_OFF = 0, KST_TILE_ACTIVE = 1, KST_TILE_GHOST = 2 };

/* Status codes. */
enum { KST_OK = 0, KST_ERR_CONFIG = -1, KST_ERR_HIP = -2, KST_ERR_HALT = -3,
       KST_ERR_NONFINITE = -4, KST_ERR_STATE = -5, KST_ERR_TOO_MANY_STEPS = -6 };

/* One flux source (RunSet%FluxSources, src/RunSettings.f90; evaluated in
 * ExplicitSourceTerms, src/Equations.f90:456-620).  time/flux/psi hold
 * n_series knots; num_cells is NumCellsInSrc (src/SetSources.f90:357-375). */
typedef struct kst_flux_source {
    double x, y, radius;
    int32_t n_series;
    int32_t num_cells;
    const double* time;
    const double* flux;
    const double* psi;
} kst_flux_source;

/* Every RunSet scalar the step reads (src/RunSettings.f90:161-340), plus the
 * derived parameters of src/Parameters.f90:629-648 (filled in by
 * kst_config_defaults / kst_config_derive). */
typedef struct kst_config {
    int32_t nx_tiles, ny_tiles;          /* nXtiles, nYtiles */
    int32_t nx_per_tile, ny_per_tile;    /* nXpertile, nYpertile */
    double dx, dy;                       /* deltaX, deltaY */
    double x_size, y_size;               /* xSize, ySize: cell centres at -size/2 + dx*(I+1/2) */
    int32_t bc;
    int32_t limiter;
    int32_t geometric_factors;           /* 1: gamma = sqrt(1+bx^2+by^2); 0: gamma = 1 */
    int32_t drag;
    int32_t erosion;
    int32_t deposition;
    int32_t ero_transition;
    int32_t morpho_damping;
    int32_t fswitch;
    int32_t morphodynamics;              /* MorphodynamicsOn */
    int32_t tile_buffer;                 /* TileBuffer */
    int32_t track_maxima;                /* UpdateMaximum* each accepted step */
    int32_t precision;                   /* KST_PREC_FP64 / KST_PREC_MIXED */
    int32_t n_sources;
    double g, rhow, rhos;
    double chezy_co, coulomb_co, manning_co;
    double pouliquen_min, pouliquen_max, pouliquen_intermediate, pouliquen_beta;
    double edwards_betastar, edwards_kappa, edwards_gamma;
    double ero_rate, ero_rate_granular, ero_depth, ero_critical_height;
    double voellmy_switch_rate, voellmy_switch_value;
    double max_pack, bed_porosity, solid_diameter, eddy_viscosity;
    double cfl, max_dt, height_threshold;
    /* derived (src/Parameters.f90:629-648) */
    double gred, rep, ws0, nsettling, critical_shields, diffusive_time_scale;
    const kst_flux_source* sources;
    /* Dirichlet boundary data (bcsHnval, bcsuval, bcsvval, bcspsival) and
     * the sponge layer's strength (SpongeStrength, default 0.2) */
    double bc_hn, bc_u, bc_v, bc_psi, sponge_strength;
} kst_config;

/* Statistics of one kst_integrate_to call. */
typedef struct kst_stats {
    int64_t steps;            /* accepted outer steps */
    int64_t rejected;         /* rolled-back attempts (refineTimeStep) */
    int64_t rhs_evals;        /* hydraulic RHS evaluations */
    int64_t cell_updates;     /* sum over accepted steps of active cells */
    int32_t active_tiles;     /* after the call */
    int32_t ghost_tiles;
    double t;                 /* simulation time after the call */
    double last_dt;           /* hydraulic dt of the last step (a Strang step spans 2 dt) */
    double min_dt, max_dt;
    int64_t redist_cells;     /* cells corrected by RedistributeCell (src/Redistribute.f90:249) */
    int64_t redist_ghost;     /* of their writes, those into ghost tiles' part of intermed3 */
} kst_stats;

typedef struct kst_handle kst_handle;

/* Fill the reference defaults (src/Parameters.f90:41-77,
 * src/SolverSettings.f90:42-51) for everything but the grid. */
void kst_config_defaults(kst_config* cfg);
/* Compute the derived parameters (src/Parameters.f90:629-648). */
void kst_config_derive(kst_config* cfg);

/* Create a handle on the current HIP device. Replaces InitialiseGrid's
 * allocation of tileContainer/intermed0..3 (src/Grid.f90:166-220). */
int kst_create(const kst_config* cfg, kst_handle** out);

/* Global vertex elevations b0, (NX+1)*(NY+1) doubles, after the tile-edge
 * equalisation of EqualiseTopographicBoundaryData
 * (src/MorphodynamicRHS.f90:588-687). Replaces GetHeights (src/dem.f90:360-393). */
int kst_set_topography(kst_handle* h, const double* b0_vertex);

/* Conserved fields w, rhoHnu, rhoHnv, Hnpsi (NX*NY each), vertex bt
 * ((NX+1)*(NY+1), may be NULL = 0; taken by the active tiles, ghost tiles
 * keep bt = 0), tile state (KST_TILE_*), time t and
 * optional initial maxima (Hnmax, may be NULL).  Ghost-tile w reconstructions
 * are rebuilt as at the end of LoadSourceConditions
 * (src/SetSources.f90:388-392). */
int kst_set_state(kst_handle* h, const double* w, const double* rhoHnu,
                  const double* rhoHnv, const double* Hnpsi, const double* bt_vertex,
                  const uint8_t* tile_state, double t, const double* hnmax0);

/* Integrate from the current time to tend with the adaptive SSP-RK3 /
 * Strang-split step of IntegrateTo (src/TimeStepper.f90:116-277). */
int kst_integrate_to(kst_handle* h, double tend, kst_stats* stats_out);

/* Take exactly n_steps accepted steps (a step that would pass tend is not
 * clamped); used for fixed-step parity runs and throughput measurement. */
int kst_step_n(kst_handle* h, int64_t n_steps, kst_stats* stats_out);

/* Download the state: w, rhoHnu, rhoHnv, Hnpsi, Hn (NX*NY each, any may be
 * NULL), vertex bt ((NX+1)*(NY+1), may be NULL; vertex (I,J) as the tile
 * holding cell (min(I,NX-1), min(J,NY-1)) sees it), tile states, time. */
int kst_get_state(kst_handle* h, double* w, double* rhoHnu, double* rhoHnv,
                  double* Hnpsi, double* Hn, double* bt_vertex, uint8_t* tile_state,
                  double* t);

/* Per-tile vertex bt, ntiles x (ny_per_tile+1) x (nx_per_tile+1): each
 * tile's own copy, as tileContainer(k)%bt holds it (src/Grid.f90:57-121);
 * copies of a shared vertex may differ in the last bits. */
int kst_get_tile_bt(kst_handle* h, double* bt_tiles);

/* Cell-centred b0, bt, db/dx, db/dy (NX*NY each, any may be NULL): fields
 * 10-13 of tileContainer(k)%u (ComputeCellCentredTopographicData,
 * src/MorphodynamicRHS.f90:308-416). */
int kst_get_cell_topography(kst_handle* h, double* b0c, double* btc, double* dbdx, double* dbdy);

/* Cell-centred bt, db/dx, db/dy (NX*NY each, layout of kst_get_cell_topography)
 * for the cells where cells[] (NX*NY) is nonzero: what LoadInitialCondition_txt
 * reads from a solution file on a restart (src/Restart.f90:216-236) and keeps,
 * since ActivateTile(restart=.true.) rebuilds only the face topography
 * (src/UpdateTiles.f90:342-384).  The cell gamma, surface-gradient prefactors,
 * the faces' CFL numerators, the activation flags and tileContainer's
 * velocities are derived again from them.  Call after kst_set_state, which
 * derives the cell topography from the vertex bt; the morphodynamic operator
 * derives it from the vertices again at its first stage, as the reference's does. */
int kst_set_cell_topography(kst_handle* h, const double* btc, const double* dbdx,
                            const double* dbdy, const uint8_t* cells);

/* Running maxima (src/TimeStepper.f90:1155-1303): 11 arrays of NX*NY doubles
 * in the order Hnmax, Hnmax_t, umax, umax_t, emax, emax_t, dmax, dmax_t,
 * psimax, psimax_t, tfirst. */
int kst_get_maxima(kst_handle* h, double* const maxima[11]);

/* Running maxima to resume from, in kst_get_maxima's order and layout (NULL
 * entries leave that array as it is): what LoadInitialCondition_txt reads
 * from the maxima file on a restart (src/Restart.f90:290-326).  Call after
 * kst_set_state, which starts the maxima afresh. */
int kst_set_maxima(kst_handle* h, const double* const maxima[11]);

/* CalculateVolume (src/Output.f90:617-735), computed on the device from the
 * resident state: out[0..5] = the Volume.txt columns 2-7 -- flow volume,
 * total bed volume, total mass, bed mass, total solids mass, bed solids
 * mass -- over the active tiles.  The reference Kahan-sums sequentially; the
 * device sums are compensated (Sum2), so both lie within an ulp or two of
 * the exact sums and may differ from each other in the last bits. */
int kst_volume(kst_handle* h, double out[6]);

/* kst_volume without allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc -- the same sums as
 * kst_volume (the same kernels and partition, so the same bits) with the
 * partial sums in a fixed device array of the library (VOL_BLOCKS_MAX = 2048
 * groups of 8 doubles plus one, 131 KB per GPU, held until the process ends)
 * instead of the handle's buffer.  The number of partial groups is capped at
 * that capacity (workgroups take further tiles in turn), so every active set
 * fits.  Calls no hipMalloc/hipFree and no allocating library routine.  The
 * handle's device must be current.  Every call on one GPU shares the array:
 * calls that overlap on one GPU (other host threads, other streams) race. */
int kst_volume_noalloc(kst_handle* h, double out[6]);

/* The per-cell columns of OutputSolutionData_txt (src/Output.f90:754-929)
 * for the active tiles, computed on the device: on return *n_tiles holds the
 * active-tile count n, tiles[0..n) the tile ids (0-based, the file writes
 * id + 1) in active-list order, and fields[(q * 14 + col) * nx_per_tile *
 * ny_per_tile + j * nx_per_tile + i] the columns 6-19 of the file -- flow
 * depth, flow speed, x and y velocity, density, solids fraction, x and y
 * flux, Hnpsi, w, base elevation, elevation change, x and y slope -- of cell
 * (i, j) of the q-th listed tile.  fields holds max_tiles * 14 * cells per
 * tile doubles and tiles max_tiles entries; KST_ERR_CONFIG if the active set
 * is larger. */
int kst_get_output(kst_handle* h, int32_t max_tiles, double* fields, int32_t* tiles, int32_t* n_tiles);

/* Copy the last error message (NUL-terminated) into buf. */
int kst_last_error(const kst_handle* h, char* buf, int buflen);

void kst_destroy(kst_handle* h);

/* Launch grids (no counterpart in the reference, which loops over its active
 * tiles on the CPU).  By default every launch covers a power-of-two cap of
 * twice the active set the library last read back (1.25 times from 256
 * tiles), and workgroups take
 * further work items in turn when the set outgrows it between read-backs.
 * kst_set_grid_cap fixes the cap at `tiles` active tiles (clamped to the
 * grid's tile count) so that grids, and the captured graphs, never change;
 * 0 restores the default.  Results do not depend on the cap. */
int kst_set_grid_cap(kst_handle* h, int tiles);

/* Autotuner (no counterpart in the reference, which loops over its tiles on
 * one CPU thread).  With enable != 0 the library times, before the first
 * launch on each new bucket of active-set sizes (active tiles rounded up to a
 * power of two), the first RK stage kernel in each candidate launch shape --
 * 16 x 16 sub-blocks on the default grid, on a grid of the blocks the device
 * holds at once (blocks then take further sub-blocks in turn) or twice that,
 * and 16 x 4 sub-blocks (FP64) -- and launches that bucket in the fastest.
 * Choices are cached for the process per grid layout, operator and bucket,
 * so later handles of the same configuration reuse them without trials.
 * The trials leave the state and the control block as they were: results
 * do not depend on the choices.  enable = 0 restores the fixed rule (16 x 4
 * below 128 sub-blocks of 16 x 16, the default grid) and forgets the
 * choices. */
int kst_autotune(kst_handle* h, int enable);
/* The choices made so far, at most max: per bucket, whether 16 x 4
 * sub-blocks were chosen, the grid limit of the 16 x 16 shape (0: default
 * grid) and the stage-1 launch time in microseconds of the four candidates
 * in the order above (0: not tried).  Returns the number written. */
int kst_autotune_choices(const kst_handle* h, int max, int32_t* bucket, int32_t* small,
                         int32_t* grid, double* us);

/* Redistribution chunks (no counterpart in the reference, whose
 * RedistributeGrid walks its sorted list one cell at a time,
 * src/Redistribute.f90:203-244).  The device walks up to `chunk` sorted
 * candidates at once (default and maximum 1024), chunk after chunk, and sorts
 * up to `lds_sort` candidates (padded to a power of two; default and maximum
 * 2048) in LDS, more in global memory.  Results do not depend on either; the
 * tests lower them to run the multi-chunk walk and the global-memory sort on
 * small candidate sets.  0 restores a default. */
int kst_set_redistribution_chunk(kst_handle* h, int chunk, int lds_sort);

/* Activation candidates (no counterpart in the reference, whose
 * CheckIfNearBoundaries walks the active list itself,
 * src/TimeStepper.f90:935-1150).  After an accepted step the device collects
 * the flagged sides whose neighbour would be activated or halt the run, and
 * walks up to `cap` of them from LDS (default and maximum 2048); a step with
 * more walks the active list in place, as the reference does.  Results do not
 * depend on it; the tests lower it to run that walk.  0 restores the
 * default. */
int kst_set_activation_candidates(kst_handle* h, int cap);

/* Diagnostics (no counterpart in the reference, which has no timers;
 * src/Messages.f90:275-283 prints only t and dt).  While profiling is on, the
 * library launches one iteration at a time and brackets every kernel launch
 * with HIP events on the handle's stream; kst_profile_read returns, per
 * kernel id below, the mean duration in ms of the launches that did work
 * (launches that only skip after a rejected check or the end of a run are
 * left out) and their counts, then resets. */
#define KST_PROFILE_KERNELS 16
enum kst_profile_kernel {
    KST_K_STAGE1 = 0,       /* hydraulic RK stages 1..4: ids 0..3 */
    KST_K_STAGE1_H2 = 4,    /* the second hydraulic operator of a Strang step: 4..7 */
    KST_K_MORPHO1 = 8,      /* morphodynamic SSP stages 1..3: 8..10 */
    KST_K_REDISTRIBUTE = 11,
    KST_K_FACE_TOPO = 12,
    KST_K_COMMIT = 13,
    KST_K_ACTIVATE = 14,
    KST_K_FINALIZE = 15
};
int kst_profile(kst_handle* h, int enable);
int kst_profile_read(kst_handle* h, double ms[KST_PROFILE_KERNELS],
                     int64_t launches[KST_PROFILE_KERNELS]);

/* ---- Batches --------------------------------------------------------------
 * No counterpart in the reference, which advances one scenario per process
 * (src/main.f90:41-133).  A batch holds up to max_slots scenario handles
 * (each created, given its topography and state as above) resident on the
 * device and advances all of them with one launch per kernel: every running
 * slot keeps its own tiles, dt, retries and time; each iteration walks one
 * work list of the active sub-blocks of all running slots.  Handles join
 * (attach) and leave (detach) between runs without rebuilding the batch's
 * hipGraph.  A batch holds 2D scenarios, hydraulic or morphodynamic, all of
 * one precision and one eddy-viscosity setting, each with its own limiter
 * (while every running scenario uses MinMod2 the kernels have it at compile
 * time); a morphodynamic scenario takes its Strang step as alone.  While attached, a handle
 * is advanced only through its batch; kst_get_state / kst_get_maxima read it
 * after kst_batch_run returns.  Results are the same bits as running the
 * handle alone. */
typedef struct kst_batch kst_batch;
enum { KST_SLOT_FREE = -1, KST_SLOT_RUNNING = 0, KST_SLOT_DONE = 1, KST_SLOT_ERROR = 2 };

int kst_batch_create(int max_slots, kst_batch** out);
/* Attach a handle to a free slot (returned in *slot). */
int kst_batch_attach(kst_batch* b, kst_handle* h, int* slot);
int kst_batch_detach(kst_batch* b, int slot);
/* Give a slot its next target: nsteps > 0 accepted steps (as kst_step_n),
 * else integrate to tend (as kst_integrate_to). */
int kst_batch_start(kst_batch* b, int slot, double tend, int64_t nsteps);
/* Advance every running slot until none is running or about max_iters
 * iterations have run; status[max_slots] receives KST_SLOT_* per slot. */
int kst_batch_run(kst_batch* b, int64_t max_iters, int* status);
/* Counters of the slot's last target (as kst_integrate_to's stats); returns
 * the scenario's error code if it stopped on one. */
int kst_batch_stats(kst_batch* b, int slot, kst_stats* out);
int kst_batch_last_error(const kst_batch* b, char* buf, int buflen);
void kst_batch_destroy(kst_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* KESTREL_STEP_H */
