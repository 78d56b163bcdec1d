/* C ABI of libfiphip: ensembles of independent fip domains on one MI355X.
 *
 * An ensemble advances many flood scenarios (each its own grid size, terrain,
 * roughness, boundary conditions, dt and target duration) together: every step is one
 * persistent stencil launch over a global list of the active tiles of all scenarios,
 * one post-step launch (one work-group per scenario: duration, dt control, BC
 * parameters) and one launch per BC stage for all scenarios; the steps run as hipGraph
 * chunks of 100. Each scenario computes exactly what a single-domain run of it computes
 * (fip.h, same mode, variant and tile rows) without sampling and save_state.
 *
 * The reference runs one domain per process (src/main.cpp:204-263); this interface has
 * no counterpart there. Return codes and fip_last_error() as in fip.h.
 */
#ifndef FIP_ENSEMBLE_H
#define FIP_ENSEMBLE_H
#include "fip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fip_ens fip_ens;

/* n scenarios (1..1024), each as for fip_create. mode, variant, tile_rows and device are
 * taken from descs[0] and apply to all; every scenario must have kst_var, or none. */
int fip_ens_create(const fip_desc* descs, int n, fip_ens** out);
int fip_ens_destroy(fip_ens* e);
/* The BC list of scenario s (fip_bc as for fip_bc_generate). */
int fip_ens_set_bcs(fip_ens* e, int s, const fip_bc* bcs, int n);
/* Target duration of every scenario (n values); resets nothing else. A scenario that has
 * reached its target stays finished; a later call moves the targets of those still
 * running. */
int fip_ens_set_run(fip_ens* e, const double* target_durations);
/* Runs up to nsteps ensemble steps, or until every scenario reached its target.
 * *active receives the number of scenarios still running. */
int fip_ens_simulate(fip_ens* e, long long nsteps, int* active, long long* steps_done);
/* Captures, without running anything, the hipGraphs that the next fip_ens_simulate(e,
 * nsteps) call launches: its steps are cut at multiples of 100 ensemble steps (each
 * chunk ends with the flood-plain kernel, which tests every scenario's iteration itself),
 * and each chunk of a length and parity is captured once and kept, as fip_prepare does
 * for one simulation. */
int fip_ens_prepare(fip_ens* e, long long nsteps);
int fip_ens_get_time(fip_ens* e, int s, double* duration, float* dt, int* iteration);
int fip_ens_download(fip_ens* e, int s, int field, int side, float* dst);
/* Launch shape of the ensemble stencil: wavefronts per work-group (1..16; 0 = the
 * default: 16 for the dense fast variant, 4 otherwise) and persistent work-groups per CU
 * (1..64), or 0 for a grid of one wavefront per active-tile entry (< 0 = default: 0).
 * Dense mode marches spans of rows with all wavefronts of a work-group in step and uses
 * no persistent work-groups. Results do not depend on it. */
int fip_ens_set_launch(fip_ens* e, int waves_per_block, int blocks_per_cu);
/* Autotuner of the ensemble (fip_tune for ensembles): times the ensemble stencil of
 * these scenarios (bcs[s] holds nbcs[s] BCs of scenario s) for every tile-row count and
 * launch shape over `steps` timed steps after as many warm-up steps, and returns the
 * fastest. Dense mode tries 2, 4, 8 and 16 wavefronts per work-group only. */
int fip_ens_tune(const fip_desc* descs, int n, const fip_bc* const* bcs, const int* nbcs,
                 int steps, fip_tune_result* out);
/* steps, tiles_total (all scenarios), tiles_computed (last step), stencil timing. */
int fip_ens_set_timing(fip_ens* e, int enable);
int fip_ens_get_stats(fip_ens* e, fip_stats* out);

#ifdef __cplusplus
}
#endif
#endif
