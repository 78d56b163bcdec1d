/* C ABI of libfiphip: the MI355X (gfx950) shallow-water time step of fip.
 *
 * Plain C types only. One simulation = one opaque handle = one HIP stream on the
 * current device. Every call returns 0 on success or a negative FIP_E* code; the message
 * of the last failure on the calling thread is returned by fip_last_error(). Nothing in
 * the library prints or exits; the fip CLI maps failures to the reference's stderr text
 * and EXIT_FAILURE.
 *
 * Fields are W x H float32, row-major and unpadded on the host side (x fastest).
 * Device storage is private: pitched with zeroed margin rows and columns.
 *
 * The entry points replace the reference's inner seam:
 *   launchComputeKernel()            src/kernels.h:3, src/kernels.cu:441-455
 *   launchFloodPlainKernel()         src/kernels.h:4, src/kernels.cu:457-463
 *   BoundaryConditionKernel::generate/launch   src/bc_kernel_gen.h:12-16
 *   controlTimestep()                src/main.cpp:25-43
 *   initDeviceBuffers()              src/simdata.cpp:54-96
 *   simulate()                       src/main.cpp:100-135 (device-resident driver)
 */
#ifndef FIP_H
#define FIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FIP_OK 0
#define FIP_E_INVALID -1   /* bad argument or configuration */
#define FIP_E_HIP -2       /* a HIP runtime call failed */
#define FIP_E_RTC -3       /* hipRTC compile of the boundary kernel failed */
#define FIP_E_STATE -4     /* call out of order (e.g. no BC kernel generated) */

/* stencil mode */
#define FIP_MODE_DENSE 0   /* every tile updated every step (ENABLE_MASK 0, kernels.cu:13) */
#define FIP_MODE_MASKED 1  /* wet/dry tile masks + compacted active-tile list */
/* arithmetic variant */
#define FIP_EXACT 0        /* IEEE div/sqrt, no FMA contraction: bit-exact with the oracle */
#define FIP_FAST 1         /* hardware rcp/sqrt/log/exp and FMA: within 1e-5 of FIP_EXACT */

/* boundary-condition type and side (src/boundary_condition.h:5-13) */
#define FIP_BC_Z 0
#define FIP_BC_Q 1
#define FIP_BC_CLOSE 2
#define FIP_BC_OPEN 3
#define FIP_SIDE_LEFT 0
#define FIP_SIDE_RIGHT 1
#define FIP_SIDE_TOP 2
#define FIP_SIDE_BOTTOM 3

/* fields for fip_download / fip_upload */
#define FIP_FIELD_H 0
#define FIP_FIELD_QX 1
#define FIP_FIELD_QY 2
#define FIP_FIELD_HMAX 3
#define FIP_FIELD_BED 4
#define FIP_FIELD_KST 5
#define FIP_SIDE_CUR 0     /* the "1" buffers (d_h1, d_qx1, d_qy1) */
#define FIP_SIDE_OTHER 1   /* the "2" buffers (d_h2, ...), dumped by save_state */

/* events returned by fip_simulate */
#define FIP_EV_NONE 0
#define FIP_EV_FINISHED 1  /* duration >= target duration */
#define FIP_EV_DUMP 2      /* a save_state time was reached; dump FIP_SIDE_OTHER + hmax */
#define FIP_EV_SAMPLES 3   /* the device sample log is nearly full: drain it, then continue */

typedef struct fip_sim fip_sim;

/* Configuration and initial state (SimData globals, src/simdata.h:23-60; copied in,
 * the caller keeps ownership of its arrays). */
typedef struct fip_desc {
  int W, H;
  float dx, dt;
  float no_data;        /* cells whose bed equals this value are never updated */
  float kst;            /* uniform Strickler coefficient when kst_var == NULL */
  const float* bed;     /* required */
  const float* h;       /* initial depth; NULL = dry */
  const float* qx;      /* NULL = 0 */
  const float* qy;      /* NULL = 0 */
  const float* kst_var; /* per-cell Strickler coefficient; NULL = uniform */
  int variable_dt;      /* 1 = CFL-controlled dt (default of the reference) */
  int staggered;        /* 1 = solver rmg (default), 0 = solver cn boundary variants */
  int mode;             /* FIP_MODE_* */
  int variant;          /* FIP_EXACT / FIP_FAST */
  int tile_rows;        /* rows per tile (16, 32 or 64); 0 = default */
  int device;           /* HIP device ordinal */
} fip_desc;

/* One boundary condition (src/boundary_condition.h, src/simdata.cpp:98-131). Either the
 * fixed values z/q or piecewise-linear series (times strictly increasing, the std::map
 * order of TimeSeries::fromCsv) are used; a series BC has zn > 0 or qn > 0. */
typedef struct fip_bc {
  int type, side, from, to;  /* cells from..to inclusive along the side */
  float z, q;
  const float* zt;
  const float* zv;
  int zn;
  const float* qt;
  const float* qv;
  int qn;
} fip_bc;

typedef struct fip_stats {
  long long steps;            /* stencil steps run */
  long long tiles_total;      /* tiles per step */
  long long tiles_computed;   /* tiles updated by the last step */
  long long stencil_launches;
  double stencil_ms;          /* summed stencil kernel time when timing is enabled */
  long long timed_launches;
  int kst_classes;            /* kSt_var read as a one-byte class per cell: the number of
                                 classes (at most 64); 0 when kSt is read as floats */
  int dense_span;             /* dense mode: rows per wavefront */
} fip_stats;

const char* fip_last_error(void);

/* initDeviceBuffers (src/simdata.cpp:54-96): allocate, upload, masks all wet, CFL 1. */
int fip_create(const fip_desc* desc, fip_sim** out);
int fip_destroy(fip_sim* sim);

/* BoundaryConditionKernel::generate (src/bc_kernel_gen.cpp:154-192): generate the
 * boundary kernel for this BC list and compile it with hipRTC for the current device. */
int fip_bc_generate(fip_sim* sim, const fip_bc* bcs, int n);
/* The generated boundary-kernel source for a BC list (what fip_bc_generate compiles),
 * for inspection and for compile checks without a GPU. Writes at most cap bytes
 * including the terminating NUL; *len receives the full length. */
int fip_bc_source(const fip_bc* bcs, int n, int W, int H, int staggered, char* buf, size_t cap,
                  size_t* len);
/* BoundaryConditionKernel::launch (src/bc_kernel_gen.cpp:194-204) with host-evaluated
 * parameters: params[2*i] = qval of BC i, params[2*i+1] = zval (from/to as generated). */
int fip_bc_launch(fip_sim* sim, const float* params);
/* Evaluate getQ/getZ (src/boundary_condition.cpp:45-77) on the device at the current
 * duration and apply the BCs: the device-resident equivalent of
 * launchBoundaryConditionKernel (src/main.cpp:85-98). */
int fip_bc_apply(fip_sim* sim);
/* The BC parameters (qval, zval) of the last device-side evaluation. */
int fip_bc_params(fip_sim* sim, float* params, int n);

/* controlTimestep (src/main.cpp:25-43) on the device, from the last step's CFL. */
int fip_control_timestep(fip_sim* sim);
/* launchComputeKernel (src/kernels.cu:441-455) + the swap of src/main.cpp:118-121:
 * one stencil step with the device dt, then duration += dt, iteration++, swap. */
int fip_compute_step(fip_sim* sim);
/* launchFloodPlainKernel (src/kernels.cu:457-463): hmax = max(hmax, h). */
int fip_floodplain(fip_sim* sim);

/* Device-resident simulate() loop (src/main.cpp:100-135, 243-246). Runs up to nsteps
 * iterations or until an event; *event receives FIP_EV_*. Sampling rows are logged on
 * the device and read with fip_read_samples. Call fip_set_run first. */
int fip_set_run(fip_sim* sim, double target_duration, int sampling_interval,
                const int* sample_x, const int* sample_y, const int* sample_type, int npoints,
                const double* save_times, int nsave);
int fip_simulate(fip_sim* sim, long long nsteps, int* event, long long* steps_done);
/* Captures, without running anything, the hipGraphs that the next fip_simulate(sim,
 * nsteps) call launches (its steps are cut at multiples of 100 iterations, the flood-plain
 * cadence of src/main.cpp:125-128, and each chunk of a length and parity is captured once
 * and kept). fip_simulate captures a chunk on first use anyway; calling this beforehand
 * takes the capture out of the first call. No counterpart in the reference, whose loop
 * launches every kernel from the host. */
int fip_prepare(fip_sim* sim, long long nsteps);
/* Drain logged samples: t[row], vals[row * npoints + p]; returns rows in *nrows. */
int fip_read_samples(fip_sim* sim, double* t, float* vals, int max_rows, int* nrows);

int fip_get_time(fip_sim* sim, double* duration, float* dt, int* iteration);
int fip_set_dt(fip_sim* sim, float dt);
int fip_download(fip_sim* sim, int field, int side, float* dst);
int fip_upload(fip_sim* sim, int field, int side, const float* src);
/* Tile geometry and the 64-bit row masks (bit k = row tile_rows*ty + k ended wet). Masks
 * are kept in masked mode; dense mode (ENABLE_MASK 0) leaves them as created. */
int fip_tiles(fip_sim* sim, int* tiles_x, int* tiles_y, int* tile_cols, int* tile_rows);
int fip_get_masks(fip_sim* sim, uint64_t* dst, int n);
int fip_get_cfl(fip_sim* sim, float* cfl_min);
/* Launch shape of the stencil: wavefronts per work-group (1..16; 0 = the default: 16 for
 * the dense fast variant, 1 for the dense exact variant, 4 in masked mode) and, in masked
 * mode, persistent work-groups per CU
 * (1..64), or 0 for a grid of one wavefront per active-tile entry (< 0 = default: 0).
 * In dense mode the wavefronts of a work-group take adjacent tile columns (the fast
 * variant marches them in step). Results do not depend on it. */
int fip_set_launch(fip_sim* sim, int waves_per_block, int blocks_per_cu);

typedef struct fip_tune_result {
  int tile_rows;         /* 16, 32 or 64 */
  int waves_per_block;
  int blocks_per_cu;     /* masked mode (0 = one wavefront per entry); -1 in dense mode */
  double stencil_ms;     /* mean stencil launch time of the chosen shape */
  int tried;             /* shapes timed */
} fip_tune_result;

/* Autotuner (no counterpart in the reference, whose tile shape is fixed at 28 x 50 by
 * src/kernels.cu:441-455): times the stencil of this configuration for every launch
 * shape (and, in masked mode, every tile-row count) over `steps` timed steps (after as
 * many warm-up steps) from the initial state, and returns the fastest. In dense mode
 * results do not depend on the choice and tile_rows is returned as given; in masked mode
 * the tile rows decide which tiles are skipped. */
int fip_tune(const fip_desc* desc, const fip_bc* bcs, int nbc, int steps, fip_tune_result* out);
int fip_set_timing(fip_sim* sim, int enable);
int fip_get_stats(fip_sim* sim, fip_stats* out);
int fip_sync(fip_sim* sim);
/* Self-check (no counterpart in the reference): the exact variant forms the reciprocal of
 * a face depth, 1/max(h, DH) (src/kernels.cu:89,116), as the correctly rounded RN(1/b)
 * from the hardware reciprocal and two Newton steps. This runs that code on every
 * binary32 b in [lo, hi] (0 < lo <= hi <= 2^64) on the given device against the IEEE
 * division and returns how many differ and the smallest that does (0 if none). */
int fip_check_rcp(int device, float lo, float hi, unsigned long long* nbad, float* first_bad);

#ifdef __cplusplus
}
#endif
#endif
