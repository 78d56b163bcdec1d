// CPU oracle for the fip shallow-water time step. TEST INFRASTRUCTURE ONLY:
// tests/, the smoke check and bench.py's cpu_baseline leg use it as the checker.
// The product (libfiphip, the fip CLI) never links or calls it.
//
// It restates the reference's hot path in IEEE host C++:
//   src/kernels.cu:23-463      stencil (rmg_kernel), mask, CFL table, flood plains
//   src/main.cpp:25-43,100-135 controlTimestep and the simulate() step order
//   src/bc_kernel_gen.cpp:11-152 boundary-condition kernel bodies
//   src/boundary_condition.cpp:6-77, src/timeseries.cpp:14-26 BC parameters
// Real type: binary32 (bit-exact target of the HIP build) or binary64.
//
// Pinning: every function is checked by tests/test_oracle.py (three stencil restatements
// that must agree, closed forms, invariants, a correctly-rounded pow sweep) and the host
// series/expression code by tests/test_cli_host.py against the reference's own
// timeseries.cpp compiled into oracle/_ref. Parity with the reference's CUDA binary
// itself (fast-math __powf, approximate division and sqrt, flush-to-zero) is unpinned:
// neither that binary nor a CUDA toolchain exists here, and the reference holds no
// golden outputs.
#pragma once
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_MODE_DENSE = 0,     // every stored cell is updated each step (ENABLE_MASK 0)
  ORC_MODE_TILED = 1,     // wet/dry tile skipping at a given tile geometry (clean)
  ORC_MODE_WARP_REF = 2,  // literal warp-lane transcription of rmg_kernel<50,false>
};

typedef struct {
  int W, H;
  float dx, dt;
  float no_data;
  float kst;          // uniform Strickler value when no kSt_var field is given
  int has_kst_var;
  int precision;      // 32 or 64
  int mode;           // ORC_MODE_*
  int tile_w;         // TILED: columns per tile; WARP_REF: lanes per warp (32)
  int tile_rows;      // TILED: rows per tile;   WARP_REF: rows per block (50)
  int warp_mask;      // WARP_REF: 1 = ENABLE_MASK semantics
  int quirks;         // WARP_REF: 1 = reproduce Q3 (first=1 warm-up store) literally
  int variable_dt;
  int staggered;      // 1 = solver rmg (default), 0 = solver cn
  int threads;        // OpenMP threads (0 = default)
  int open_pitch;     // 0: the top/bottom open-BC sign is +-1 (this build); > 0: it is
                      // +-open_pitch, the reference's (float)(idx - idx_neighbor) of a row
                      // pitch in elements (quirk Q8, src/bc_kernel_gen.cpp:132)
} orc_desc;

typedef struct orc_state orc_state;

orc_state* orc_create(const orc_desc* d, const float* bed, const float* h, const float* qx,
                      const float* qy, const float* kst_var);
void orc_destroy(orc_state* s);

// Boundary conditions, applied in list order. type: 0=z 1=q 2=close 3=open;
// side: 0=left 1=right 2=top 3=bottom. Series may be given for z and/or q
// (n points of time/value, already in std::map order: strictly increasing time).
int orc_add_bc(orc_state* s, int type, int side, int from, int to, float z, float q,
               const float* zt, const float* zv, int zn, const float* qt, const float* qv, int qn);

// One simulate() iteration (src/main.cpp:100-135) without sampling and dumps.
void orc_simulate_step(orc_state* s);
// Pieces of simulate(), for tests that drive the order themselves.
void orc_control_timestep(orc_state* s);
void orc_apply_bcs(orc_state* s);            // BC(duration) on the current buffers
void orc_stencil(orc_state* s);              // one rmg step, duration += dt, swap
void orc_floodplain(orc_state* s);
// BC parameter evaluation at the current duration (qval, zval for BC i).
void orc_bc_params(orc_state* s, int i, float* qval, float* zval);
float orc_sum_water_depths(orc_state* s, int i, float z);
float orc_interp(const float* t, const float* v, int n, float time);
float orc_pow233(float x);
// Exhaustive check of fip_pow233 on [lo, hi] against long double pow (see fip_oracle.cpp).
void orc_pow233_sweep(float lo, float hi, long long* n, long long* nbad, float* bad_x, int cap);

void orc_get_time(orc_state* s, double* t, double* dt, int* iteration);
void orc_set_dt(orc_state* s, double dt);
// field: 0=h 1=qx 2=qy 3=hmax 4=bed 5=kst; side: 0 = current ("1" buffers), 1 = other
void orc_get_field(orc_state* s, int field, int side, float* out);
void orc_get_field64(orc_state* s, int field, int side, double* out);
void orc_set_field(orc_state* s, int field, int side, const float* in);
// Tile masks and the CFL minimum written by the last stencil step.
int orc_num_tiles(orc_state* s);
void orc_get_masks(orc_state* s, uint64_t* out);
double orc_cfl_min(orc_state* s);
long long orc_cells_updated(orc_state* s);   // stored cells written by the last step
long long orc_tiles_computed(orc_state* s);

#ifdef __cplusplus
}
#endif
