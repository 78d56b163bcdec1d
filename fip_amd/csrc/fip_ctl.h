// Device-resident control block of one simulation: time, dt, iteration, ping-pong
// side, CFL accumulators, the active-tile schedule and the host-event flags. It lives in
// device memory and is shared by the precompiled kernels and the run-time compiled
// boundary-condition kernel (this file is also handed to hipRTC as a header).
#pragma once

#define FIP_MAX_BC 256
#define FIP_MAX_SAMPLE_POINTS 64
#define FIP_SAMPLE_LOG_ROWS 256
#define FIP_MAX_SAVE 256
#ifndef FIP_KST_CLASSES
#define FIP_KST_CLASSES 64
#endif

// Stop requests raised by the post-step kernel (bits).
#define FIP_HALT_FINISHED 1  // duration >= target: the loop of src/main.cpp:243 ends
#define FIP_HALT_DUMP 2      // a save_state time was reached: the host dumps, then resumes

typedef struct fip_ctl {
  double duration;       // simulated time (src/simdata.cpp:19), double as in the reference
  double target;         // target_duration
  double sample_timer;   // Sampling::timer
  double save_times[FIP_MAX_SAVE];  // save_state times, sorted (src/simdata.cpp:304-309)
  float dt;              // current timestep
  float cfl_prev;        // CFL minimum written by the last stencil step (starts at 1)
  unsigned cfl_acc[2];   // float bits; stencil of iteration n min-reduces into [n & 1]
  int iteration;         // the current ("1") buffers are index iteration & 1
  // Stop requests. The post-step kernel of a step with parity p = iteration & 1 writes
  // halt_req[p]; every kernel of the next step (parity p^1) reads halt_req[p] and does
  // nothing when it is set, and a skipped post kernel copies the request forward. No
  // kernel reads a word that another block of the same launch writes.
  int halt_req[2];
  int variable_dt;
  int sampling_interval; // seconds; 0 = off
  int nsave;             // number of valid entries in save_times
  int save_pos;          // next entry of save_times
  int nsample_rows;      // rows written to the sample log since the host last drained it
  int nsample_points;
  int sample_overflow;
  int pre_sampled;       // the pre-first-step kernel logged the sample of the first call
  // active-tile schedule (masked mode), double-buffered by iteration parity
  unsigned list_count[2];
  unsigned work_next[2];   // reset each step; the list kernels stride and do not use it
  // boundary-condition parameters of the last BC application
  float bc_q[FIP_MAX_BC];
  float bc_z[FIP_MAX_BC];
  float swd_z[FIP_MAX_BC];    // sumWaterDepths cache: last z and its value
  float swd_val[FIP_MAX_BC];
  int swd_valid[FIP_MAX_BC];
  // sample points and the log of sampled rows
  int sp_x[FIP_MAX_SAMPLE_POINTS];
  int sp_y[FIP_MAX_SAMPLE_POINTS];
  int sp_type[FIP_MAX_SAMPLE_POINTS];  // 0=z 1=h 2=qx 3=qy
  double sample_t[FIP_SAMPLE_LOG_ROWS];
  float sample_v[FIP_SAMPLE_LOG_ROWS * FIP_MAX_SAMPLE_POINTS];
} fip_ctl;

// Pointers and sizes of one simulation, passed by value to every kernel.
typedef struct fip_grid {
  float* h[2];
  float* qx[2];
  float* qy[2];
  float* hmax;
  const float* bed;
  const float* kst;       // NULL when kSt is uniform
  // kSt by class: when kSt_var takes at most FIP_KST_CLASSES distinct values, kidx holds
  // the class of every cell (same element layout as the float fields, one byte per
  // element) and ktab the class values; NULL otherwise
  const unsigned char* kidx;
  const float* ktab;
  unsigned long long* mask;   // one 64-bit row mask per tile
  unsigned* list[2];          // active tiles: (tile << 12) | (lo << 6) | hi
  const float* series;        // BC time-series tables
  fip_ctl* ctl;
  int W, H, P;            // P = row stride in floats; element (i, j) is at [j * P + i]
  int margin;             // elements from the start of an allocation to element (0, 0)
  int field_bytes;        // bytes per field allocation (buffer-resource range)
  int tiles_x, tiles_y, tile_rows;
  float dx, no_data, kst_uniform;
} fip_grid;

// ---- ensembles: many independent domains advanced by one launch per kernel and step ----
#define FIP_ENS_MAX 1024

// One scenario of an ensemble: its domain, its BC list [bc0, bc0 + nbc) and the index of
// its first tile in the ensemble's global tile numbering.
typedef struct fip_ens_scn {
  fip_grid g;
  int bc0, nbc;
  int tile_base;
  int staggered;
} fip_ens_scn;

// A boundary condition of an ensemble scenario (series offsets index the ensemble's
// series table).
typedef struct fip_ens_bc {
  int type, side, from, to;
  int zoff, zn, qoff, qn;
  float z, q;
} fip_ens_bc;

// Up to 256 consecutive border positions of one BC, applied by one work-group.
typedef struct fip_ens_chunk {
  int scn, bc, pos0, n;
} fip_ens_chunk;

// Ensemble-wide schedule: the global active-tile list of each parity, entries
// (scenario << 40) | (tile << 12) | (lo << 6) | hi.
typedef struct fip_ens_ctl {
  unsigned list_count[2];
  unsigned work_next[2];   // reset each step; the list kernels stride and do not use it
  int finished[FIP_ENS_MAX];  // set by the post kernel when a scenario reached its target
} fip_ens_ctl;
