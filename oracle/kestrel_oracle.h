/*
 * kestrel_oracle.h -- CPU restatement of Kestrel's finite-volume step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is checked
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker or as the timed CPU baseline; the
 * product (kestrel_amd/) never calls it.
 *
 * It keeps the reference's data model -- tiles with AoS-by-field arrays, a
 * tileContainer plus intermed0..3 deep copies, sorted active-tile list and
 * ghost tiles with frozen data -- so that every sweep can be compared with
 * the reference routine it restates (file:line citations in the .c file).
 */
#ifndef KESTREL_ORACLE_H
#define KESTREL_ORACLE_H

#include <stdint.h>
#include "../include/kestrel_step.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ko_sim ko_sim;

ko_sim* ko_create(const kst_config* cfg);
void ko_destroy(ko_sim* s);
const char* ko_error(const ko_sim* s);

/* Global vertex b0, (NX+1)*(NY+1), row-major (same layout as kst_set_topography). */
int ko_set_topography(ko_sim* s, const double* b0_vertex);

/* Initial conditions (src/SetSources.f90:47-392): caps (shape 0 flat,
 * 1 para, 2 level), flux-source cell counting, then the closing
 * ReconstructwAtEdges sweep over ghost tiles. */
int ko_add_caps(ko_sim* s, int ncaps, const double* x, const double* y, const double* radius,
                const double* height, const double* psi, const double* u, const double* v,
                const int* shape);
/* Counts NumCellsInSrc for cfg->sources (writes into num_cells_out[n_sources])
 * and activates the tiles holding source cells. */
int ko_add_flux_sources(ko_sim* s, int32_t* num_cells_out);
int ko_finish_sources(ko_sim* s);

/* Direct state load (the C-ABI's kst_set_state semantics). */
int ko_set_state(ko_sim* s, const double* w, const double* rhoHnu, const double* rhoHnv,
                 const double* Hnpsi, const double* bt_vertex, const uint8_t* tile_state,
                 double t, const double* hnmax0);

int ko_integrate_to(ko_sim* s, double tend, kst_stats* st);
int ko_step_n(ko_sim* s, int64_t n, kst_stats* st);

int ko_get_state(ko_sim* s, double* w, double* rhoHnu, double* rhoHnv, double* Hnpsi,
                 double* Hn, double* bt_vertex, uint8_t* tile_state, double* t);
int ko_get_maxima(ko_sim* s, double* const maxima[11]);

/* Full tile-level dump for cross-checks: all 13 fields of every allocated
 * tile's u (tileContainer), cell-major global layout, NX*NY per field
 * (unallocated tiles give 0). */
int ko_set_maxima(ko_sim* s, const double* const mx[11]);
int ko_set_cell_topography(ko_sim* s, const double* btc, const double* dbdx, const double* dbdy,
                           const uint8_t* cells);
int ko_get_fields13(ko_sim* s, double* out13);
/* CalculateVolume (src/Output.f90:617-735): the Volume.txt columns 2-7
 * (volume, total bed volume, total mass, bed mass, total solids mass, bed
 * solids mass) of tileContainer's active tiles. */
int ko_volume(ko_sim* s, double out[6]);
/* Its per-cell terms (flow volume, mass, bed, solids) in summation order;
 * returns the cell count, or -1 past cap cells. */
int ko_volume_terms(ko_sim* s, double* terms, int64_t cap);
/* Ghost-tile w slopes (uLimX/uLimY of field w), global NX*NY each. */
int ko_get_ghost_slopes(ko_sim* s, double* sx, double* sy);

/* One CalculateHydraulicRHS on tileContainer (substep 1 semantics without
 * touching grid%dt): writes ddtExplicit (4*NX*NY) and ddtImplicit (4*NX*NY)
 * in global layout and returns the advised time step. */
int ko_rhs(ko_sim* s, int substep, double t, double* ddte, double* ddti, double* advised);

/* Number of active / ghost tiles and the sorted active list (0-based ids). */
int ko_tile_lists(ko_sim* s, int32_t* active, int32_t* nactive, int32_t* ghost, int32_t* nghost);

/* Per-tile vertex bt of tileContainer: ntiles x (ny+1) x (nx+1), the
 * reference's tileContainer(k)%bt (shared vertices may differ by tile). */
int ko_get_tile_bt(ko_sim* s, double* bt_tiles);

/* CalculateMorphodynamicRHS on a copy of tileContainer: EminusD (NX*NY,
 * global) and ddtExplicitBt per tile (ntiles x (ny+1) x (nx+1)). */
int ko_morpho_rhs(ko_sim* s, double* eminusd, double* dtbt_tiles);

/* Probe of the first redistribution of a run: after ko_probe_enable, the
 * first morphodynamic step that redistributes (with on = 2: the first
 * morphodynamic step that is not refined) records intermed0, intermed3
 * before and after RedistributeGrid (13 fields, global, and per-tile vertex
 * bt) and the candidate cells in the order the check loop marked them. */
void ko_probe_enable(ko_sim* s, int on);   /* on = 2: first morphodynamic step */
double ko_probe_dt(ko_sim* s);
int ko_probe_count(ko_sim* s);
int ko_probe_get(ko_sim* s, int which, double* f13, double* bt_tiles);
int ko_probe_candidates(ko_sim* s, int32_t* tile, int32_t* i, int32_t* j, double* excess,
                        uint8_t* tile_state);

/* Set the per-tile vertex bt of every allocated tile, then recompute the
 * cell (all allocated tiles) and face (active tiles) topography. */
int ko_set_tile_bt(ko_sim* s, const double* bt_tiles);

/* Overwrite tileContainer's fields 1..13 (global layout) and per-tile vertex
 * bt (may be NULL) of every allocated tile. */
int ko_set_fields13(ko_sim* s, const double* f13, const double* bt_tiles);

/* AddToRedistList over the given candidates (1-based) and RedistributeGrid
 * with intermed0 = (f0, bt0), intermed3 = (f3, bt3): intermed3 afterwards. */
int ko_redistribute(ko_sim* s, const double* f0, const double* bt0, const double* f3,
                    const double* bt3, int n, const int32_t* tile, const int32_t* i,
                    const int32_t* j, const double* excess, double* fout, double* btout,
                    int* failed);

/* Record (t, dt) pairs of the next `cap` accepted steps into buf (the
 * caller keeps buf alive; cap = 0 stops); returns how many the previous
 * buffer received. */
int64_t ko_step_log(ko_sim* s, double* buf, int64_t cap);

/* lim(a, b) of limiter kind KST_LIMITER_* (src/Limiters.f90:83-184). */
double ko_limiter(int kind, double a, double b);

#ifdef __cplusplus
}
#endif

#endif
