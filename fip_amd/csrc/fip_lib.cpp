// libfiphip: host side of the C ABI (include/fip.h). Allocation and upload, the hipRTC
// boundary-kernel generator, the fine-grained step seam and the device-resident
// simulate() loop with hipGraph-captured chunks.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/fip.h"
#include "fip_ctl.h"
#include "fip_embed.h"  // generated: the device headers as strings for hipRTC
#include "fip_internal.h"

hipError_t fip_launch_stencil_exact(const fip_grid& g, int par, int masked, int blocks, int wpb, int span,
                                    int flood, hipStream_t st);
hipError_t fip_launch_stencil_fast(const fip_grid& g, int par, int masked, int blocks, int wpb, int span,
                                   int flood, hipStream_t st);
hipError_t fip_stencil_dense_waves_per_cu_exact(int kvar, int wpb, int* waves);
hipError_t fip_stencil_dense_waves_per_cu_fast(int kvar, int wpb, int* waves);
hipError_t fip_launch_misc(int which, const fip_grid& g, int arg, int blocks, hipStream_t st);
hipError_t fip_launch_check_rcp(unsigned lo, unsigned n, unsigned long long* nbad, unsigned* first,
                                hipStream_t st);

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPTRY(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(FIP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
  } while (0)

constexpr int kTileCols = 60;
constexpr int kMarginCols = 64;
constexpr int kMarginRows = 8;
constexpr int kStepsPerGraph = 100;  // a multiple of 2 and the flood-plain cadence

struct BCDesc {
  int type, side, from, to;
  float z, q;
  int zoff, zn, qoff, qn;
};

std::string hexf(float v) {
  if (std::isnan(v)) return "__builtin_nanf(\"\")";
  if (std::isinf(v)) return v > 0 ? "__builtin_inff()" : "(-__builtin_inff())";
  char buf[64];
  std::snprintf(buf, sizeof buf, "%af", (double)v);
  return buf;
}

}  // namespace

struct fip_sim {
  fip_desc d{};
  int dev = 0;
  hipStream_t stream = nullptr;
  fip_grid g{};
  float* base_alloc[9] = {};
  unsigned char* kidx_alloc = nullptr;  // kSt classes (see fip_grid::kidx)
  float* ktab_alloc = nullptr;
  size_t field_elems = 0;  // elements per allocation (with margins)
  long interior_off = 0;   // offset of element (0,0)
  int ntiles = 0;
  int host_iter = 0;
  bool masked = false;
  bool fast = false;
  int cus = 256;           // compute units of the device
  int wpb = 4;             // wavefronts per stencil work-group (see default_wpb)
  int bpc = 0;             // masked mode: persistent work-groups per CU (0 = one wave per entry)
  int stencil_blocks = 0;
  int sched_blocks = 0;
  int dense_span = 0;      // dense mode: rows per wavefront
  int dense_blocks = 0;    // dense mode: work-groups per step
  // boundary kernel
  hipModule_t module = nullptr;
  hipFunction_t bc_post_fn = nullptr;
  hipFunction_t bc_apply_fn = nullptr;
  std::vector<int> bc_stage_items;  // border positions per apply stage
  int nbc = 0;
  bool bc_ready = false;
  float* d_series = nullptr;
  // run state
  bool started = false;
  bool run_set = false;
  int npoints = 0;
  // captured chunks: graph of `steps` steps starting at parity `par` (plus the flood-plain
  // kernel, which tests iteration % 100 itself), most recently built last
  struct Chunk {
    int par, steps;
    bool flood_first;
    hipGraphExec_t exec;
  };
  std::vector<Chunk> graphs;
  bool graph_failed = false;
  long long flood_done_iter = -1;  // dense mode: the flood-plain iteration last taken by the separate pass
  fip_internal::StepHook post_hook;  // extra work after each step of the device-resident loop
  // stats / timing
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  fip_stats stats{};
  std::string arch;
  int* pin = nullptr;  // pinned host copy of the control words read after every chunk
};

namespace {

// Dense mode: one wavefront per span of rows of a tile column, the wpb wavefronts of a
// work-group on adjacent tile columns of one span row (the last group of a span row may
// hold idle wavefronts past the grid). Interior spans march span + 2 rows (two warm-up
// rows); span + 2 is kept a multiple of the 6-row register ring.
// * Fast variant: spans sized so that the whole grid is resident at once (every SIMD
//   holds its full share of equally long wavefronts: no tail of a few wavefronts running
//   alone). Its cost per row is the same wet or dry.
// * Exact variant: 16-row spans, many more wavefronts than slots. Its cost per row
//   depends on the water (the friction power and divisions run only where there is flow),
//   so with one span per slot the step lasted as long as the wettest spans; with short
//   spans the dispatcher hands the next span to whichever slot frees first. At
//   5667 x 6500, 30 % wet: 124-row spans 0.309 ms, 46 rows 0.298, 22 rows 0.279, 16 rows
//   0.277, 10 rows 0.295 (the warm-up rows' share grows).
constexpr int kExactDenseSpan = 16;
int plan_dense(fip_sim* s) {
  const int kv = s->g.kidx ? 2 : s->g.kst ? 1 : 0;
  int wpc = 0;
  hipError_t e = s->fast ? fip_stencil_dense_waves_per_cu_fast(kv, s->wpb, &wpc)
                         : fip_stencil_dense_waves_per_cu_exact(kv, s->wpb, &wpc);
  if (e != hipSuccess) return fail(FIP_E_HIP, std::string("occupancy query: ") + hipGetErrorString(e));
  if (wpc < 1) return fail(FIP_E_HIP, "the dense stencil kernel cannot be resident");
  const long slots = (long)s->cus * wpc;
  const int tx = s->g.tiles_x, H = s->g.H;
  const long per_col = std::max<long>(1, slots / ((long)(tx + s->wpb - 1) / s->wpb * s->wpb));
  int span = (int)((H + per_col - 1) / per_col);
  span = s->fast ? std::max(span, 16) : kExactDenseSpan;
  while ((span + 2) % 6) span++;
  const int nspans = (H + span - 1) / span;
  s->dense_span = span;
  s->dense_blocks = nspans * ((tx + s->wpb - 1) / s->wpb);
  return FIP_OK;
}

// Wavefronts per work-group by default: 16 for the dense fast variant (its work-groups
// march in step; 16 measured 0.5 % faster than 8 at 5667 x 6500), 4 otherwise.
// Dense fast: 16 wavefronts marching in step; dense exact: one wavefront per work-group
// (its arithmetic, not the row order, bounds it; measured 2 % faster than 4 at 5667 x 6500);
// masked: 4.
int default_wpb(const fip_sim* s) { return s->masked ? 4 : s->fast ? 16 : 1; }

// flood (dense mode): the step also takes the flood plains of the state it reads.
int launch_stencil(fip_sim* s, int par, int flood = 0) {
  const int kind = s->masked ? (s->bpc > 0 ? 2 : 1) : 0;
  const int blocks = kind == 2 ? s->stencil_blocks
                     : kind == 1 ? (s->ntiles + s->wpb - 1) / s->wpb : s->dense_blocks;
  hipError_t e;
  if (s->timing) HIPTRY(hipEventRecord(s->ev0, s->stream));
  e = s->fast ? fip_launch_stencil_fast(s->g, par, kind, blocks, s->wpb, s->dense_span, flood, s->stream)
              : fip_launch_stencil_exact(s->g, par, kind, blocks, s->wpb, s->dense_span, flood, s->stream);
  if (e != hipSuccess) return fail(FIP_E_HIP, std::string("stencil launch: ") + hipGetErrorString(e));
  if (s->timing) {
    HIPTRY(hipEventRecord(s->ev1, s->stream));
    HIPTRY(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIPTRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->stats.stencil_ms += ms;
    s->stats.timed_launches++;
  }
  s->stats.stencil_launches++;
  return FIP_OK;
}

int launch_misc(fip_sim* s, int which, int arg, int blocks) {
  hipError_t e = fip_launch_misc(which, s->g, arg, blocks, s->stream);
  if (e != hipSuccess) return fail(FIP_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return FIP_OK;
}

// Boundary kernel modes.
enum { BC_PRE0 = 0, BC_POST = 1, BC_APPLY_DEVICE = 2, BC_APPLY_HOST = 3 };

int launch_bc(fip_sim* s, int mode, int par) {
  if (!s->bc_ready) return fail(FIP_E_STATE, "boundary kernel not generated (call fip_bc_generate)");
  if (mode != BC_APPLY_HOST) {
    void* args[] = {&s->g, &mode, &par};
    HIPTRY(hipModuleLaunchKernel(s->bc_post_fn, 1, 1, 1, 256, 1, 1, 0, s->stream, args, nullptr));
  }
  const int sched = (s->masked && (mode == BC_PRE0 || mode == BC_POST)) ? s->sched_blocks : 0;
  const int nst = std::max<int>(1, (int)s->bc_stage_items.size());
  for (int st = 0; st < nst; st++) {
    const int items = st < (int)s->bc_stage_items.size() ? s->bc_stage_items[st] : 0;
    int napply = (items + 255) / 256;
    const int blocks = napply + (st == 0 ? sched : 0);
    if (blocks == 0) continue;
    void* args[] = {&s->g, &mode, &par, (void*)&st, &napply};
    HIPTRY(hipModuleLaunchKernel(s->bc_apply_fn, blocks, 1, 1, 256, 1, 1, 0, s->stream, args, nullptr));
  }
  return FIP_OK;
}

template <class T>
int ctl_write(fip_sim* s, size_t off, const T* v, size_t n = 1) {
  HIPTRY(hipMemcpyAsync((char*)s->g.ctl + off, v, sizeof(T) * n, hipMemcpyHostToDevice, s->stream));
  HIPTRY(hipStreamSynchronize(s->stream));
  return FIP_OK;
}

template <class T>
int ctl_read(fip_sim* s, size_t off, T* v, size_t n = 1) {
  HIPTRY(hipMemcpyAsync(v, (char*)s->g.ctl + off, sizeof(T) * n, hipMemcpyDeviceToHost, s->stream));
  HIPTRY(hipStreamSynchronize(s->stream));
  return FIP_OK;
}

#define CTL_OFF(f) offsetof(fip_ctl, f)

// The control words fip_simulate reads after every chunk, from `iteration` through
// `nsample_rows`, in one copy to pinned memory.
struct CtlStatus {
  int iteration, halt_req[2], nsample_rows;
};
constexpr size_t kStatusWords = (CTL_OFF(nsample_rows) - CTL_OFF(iteration)) / sizeof(int) + 1;
static_assert(CTL_OFF(halt_req) == CTL_OFF(iteration) + sizeof(int), "control block layout");
static_assert(kStatusWords <= 16, "control block layout");

int ctl_status(fip_sim* s, CtlStatus* st) {
  HIPTRY(hipMemcpyAsync(s->pin, (char*)s->g.ctl + CTL_OFF(iteration), kStatusWords * sizeof(int),
                        hipMemcpyDeviceToHost, s->stream));
  HIPTRY(hipStreamSynchronize(s->stream));
  st->iteration = s->pin[0];
  st->halt_req[0] = s->pin[1];
  st->halt_req[1] = s->pin[2];
  st->nsample_rows = s->pin[kStatusWords - 1];
  return FIP_OK;
}

// Boundary kernels for one BC list: the counterpart of the string templates of
// src/bc_kernel_gen.cpp:11-152, split in two launches.
//
// fip_bc_post - one work-group: the serial per-step control. Modes:
//   PRE0  - before the first step: controlTimestep, sample, BC parameters at t
//           (src/main.cpp:102-114)
//   POST  - after a stencil step of parity par: the sample of the simulate() call that
//           step belongs to (taken now; the stencil does not write its input side),
//           duration += dt, iteration++, BC parameters at the new t, save_state / end of
//           run requests, controlTimestep of the next call (src/main.cpp:100-135)
//   APPLY_DEVICE - BC parameters at t only
// fip_bc_apply - many work-groups, one border position per thread: applies the BCs with
//   the parameters in the control block (APPLY_HOST: uploaded by the host, as
//   BoundaryConditionKernel::launch). The reference applies the list in order within one
//   thread per position; BC j must run after every earlier BC whose border strip (the two
//   cells deep at each position it reads or writes) intersects its own, so the list is cut
//   into stages of pairwise-disjoint BCs, one launch each. Extra work-groups of the first
//   stage build the next active-tile list (masked mode, PRE0/POST).
// stage_items receives the number of border positions of each stage.
std::string bc_source(const std::vector<BCDesc>& bcs, bool staggered, int W, int H,
                      std::vector<int>* stage_items) {
  std::string s;
  s += "#include \"fip_ctl.h\"\n#include \"fip_detmath.h\"\n#include \"fip_sched.h\"\n"
       "#include \"fip_step_common.h\"\n#include \"fip_bc_common.h\"\n\n";
  char buf[1024];
  const int n = (int)bcs.size();
  const int nb = std::max(1, n);
  int nser = 0;
  for (const BCDesc& b : bcs) nser = std::max(nser, std::max(b.zoff + 2 * b.zn, b.qoff + 2 * b.qn));
  const bool ser_lds = nser > 0 && nser <= 8192;
  bool anyq = false;
  for (const BCDesc& b : bcs) anyq |= b.type == FIP_BC_Q;

  // ---- fip_bc_post ----
  s += "extern \"C\" __global__ __launch_bounds__(256) void fip_bc_post(fip_grid g, int mode, int par) {\n"
       "  fip_ctl* c = g.ctl;\n"
       "  const int tid = threadIdx.x;\n";
  std::snprintf(buf, sizeof buf,
                "  __shared__ int s_skip, s_row, s_np, s_cur;\n"
                "  __shared__ int s_need[%d];\n  __shared__ float s_zv[%d], s_swd[%d];\n",
                nb, nb, nb);
  s += buf;
  if (anyq) s += "  __shared__ float s_terms[256];\n";
  if (ser_lds) {
    std::snprintf(buf, sizeof buf,
                  "  __shared__ float s_series[%d];\n"
                  "  for (int i = tid; i < %d; i += blockDim.x) s_series[i] = g.series[i];\n"
                  "  const float* ser = s_series;\n",
                  nser, nser);
    s += buf;
  } else {
    s += "  const float* ser = g.series;\n";
  }
  s += "  (void)ser;\n"
       "  double dur = 0.0, target = 0.0, timer = 0.0, next_save = 0.0;\n"
       "  float dt = 0.f, cflp = 0.f;\n"
       "  unsigned acc = 0u;\n"
       "  int it = 0, vdt = 0, interval = 0, spos = 0, nsave = 0, hreq = 0, pre = 0, nrows = 0, np = 0;\n";
  for (int i = 0; i < n; i++)
    if (bcs[i].type == FIP_BC_Q) {
      std::snprintf(buf, sizeof buf, "  int swd_ok%d = 0; float swd_z%d = 0.f, swd_v%d = 0.f;\n", i, i, i);
      s += buf;
    }
  // one batch of loads: no store of this thread precedes them
  s += "  if (tid == 0) {\n"
       "    hreq = c->halt_req[par ^ 1]; dur = c->duration; dt = c->dt; it = c->iteration;\n"
       "    acc = c->cfl_acc[par]; cflp = c->cfl_prev; vdt = c->variable_dt; target = c->target;\n"
       "    timer = c->sample_timer; interval = c->sampling_interval; spos = c->save_pos;\n"
       "    nsave = c->nsave; next_save = c->save_times[spos < FIP_MAX_SAVE ? spos : 0];\n"
       "    pre = c->pre_sampled; nrows = c->nsample_rows; np = c->nsample_points;\n";
  for (int i = 0; i < n; i++)
    if (bcs[i].type == FIP_BC_Q) {
      std::snprintf(buf, sizeof buf,
                    "    swd_ok%d = c->swd_valid[%d]; swd_z%d = c->swd_z[%d]; swd_v%d = c->swd_val[%d];\n",
                    i, i, i, i, i, i);
      s += buf;
    }
  s += "  }\n  __syncthreads();\n"
       "  if (tid == 0) {\n"
       "    int skip = 0, row = -1, cur = par;\n"
       "    if (mode == 1) {\n"
       "      if (hreq) { c->halt_req[par] = hreq; skip = 1; }\n"
       "      else {\n"
       "        if (pre) c->pre_sampled = 0;\n"
       "        else row = fip_sample_begin(c, dur, timer, interval, nrows);\n"
       "        dur = dur + (double)dt;\n"
       "        it = it + 1;\n"
       "        cflp = fip_u2f(acc);\n"
       "        c->duration = dur; c->iteration = it; c->cfl_prev = cflp;\n"
       "        c->cfl_acc[par] = fip_f2u(1.0f); c->list_count[par] = 0u; c->work_next[par] = 0u;\n"
       "      }\n"
       "    } else if (mode == 0) {\n"
       "      if (vdt) { dt = fip_control_dt(dt, cflp); c->dt = dt; }\n"
       "      row = fip_sample_begin(c, dur, timer, interval, nrows);\n"
       "      c->pre_sampled = 1;\n"
       "    }\n"
       "    if (!skip) {\n"
       "      const float tf = (float)dur;\n"
       "      (void)tf;\n";
  // getZ of each BC at the new time (src/boundary_condition.cpp:66-77); a series BC
  // without a z series reads z as 0
  for (int i = 0; i < n; i++) {
    const BCDesc& b = bcs[i];
    const bool series = b.zn > 0 || b.qn > 0;
    std::string zexpr;
    if (!series) {
      zexpr = hexf(b.z);
    } else if (b.zn > 0) {
      std::snprintf(buf, sizeof buf, "fip_interp(ser + %d, ser + %d, %d, tf)", b.zoff, b.zoff + b.zn, b.zn);
      zexpr = buf;
    } else {
      zexpr = "0.f";
    }
    std::snprintf(buf, sizeof buf, "      s_zv[%d] = %s;\n", i, zexpr.c_str());
    s += buf;
    if (b.type == FIP_BC_Q) {
      std::snprintf(buf, sizeof buf,
                    "      s_need[%d] = !(swd_ok%d && fip_f2u(swd_z%d) == fip_f2u(s_zv[%d]));\n", i, i,
                    i, i);
      s += buf;
    }
  }
  s += "    }\n"
       "    s_skip = skip; s_row = row; s_np = np; s_cur = cur;\n"
       "  }\n  __syncthreads();\n"
       "  if (s_skip) return;\n"
       "  if (s_row >= 0 && tid < s_np) fip_sample_point(g, s_cur, s_row, tid);\n"
       "  if (mode == 3) return;\n";
  // BoundaryCondition::sumWaterDepths (src/boundary_condition.cpp:6-18) when z changed:
  // the terms are formed in parallel, thread 0 adds them in the reference's order
  for (int i = 0; i < n; i++) {
    const BCDesc& b = bcs[i];
    if (b.type != FIP_BC_Q) continue;
    const char* idx = b.side == FIP_SIDE_LEFT    ? "(long)k * g.P"
                      : b.side == FIP_SIDE_RIGHT ? "(long)k * g.P + (g.W - 1)"
                      : b.side == FIP_SIDE_TOP   ? "(long)(g.H - 1) * g.P + k"
                                                 : "(long)k";
    std::snprintf(buf, sizeof buf,
                  "  if (s_need[%d]) {\n"
                  "    const float zv = s_zv[%d];\n"
                  "    float ah_sum = 0.f;\n"
                  "    for (int base = %d; base < %d; base += 256) {\n"
                  "      const int k = base + tid;\n"
                  "      if (k < %d) s_terms[tid] = fip_swd_term(zv, g.bed[%s]);\n"
                  "      __syncthreads();\n"
                  "      if (tid == 0) {\n"
                  "        const int m = %d - base < 256 ? %d - base : 256;\n"
                  "        for (int e = 0; e < m; e++) ah_sum += s_terms[e];\n"
                  "      }\n"
                  "      __syncthreads();\n"
                  "    }\n"
                  "    if (tid == 0) s_swd[%d] = ah_sum * g.dx;\n"
                  "  }\n",
                  i, i, b.from, b.to, b.to, idx, b.to, b.to, i);
    s += buf;
  }
  s += "  if (tid != 0) return;\n";
  for (int i = 0; i < n; i++) {
    const BCDesc& b = bcs[i];
    const bool series = b.zn > 0 || b.qn > 0;
    std::string qexpr;
    if (b.type == FIP_BC_Q) {
      std::snprintf(buf, sizeof buf,
                    "  {\n    float swd = swd_v%d;\n"
                    "    if (s_need[%d]) { swd = s_swd[%d]; c->swd_z[%d] = s_zv[%d]; c->swd_val[%d] = swd; c->swd_valid[%d] = 1; }\n",
                    i, i, i, i, i, i, i);
      s += buf;
      std::string qs;
      if (!series) {
        qs = hexf(b.q);
      } else if (b.qn > 0) {
        std::snprintf(buf, sizeof buf, "fip_interp(ser + %d, ser + %d, %d, (float)dur)", b.qoff,
                      b.qoff + b.qn, b.qn);
        qs = buf;
      } else {
        qs = "0.f";
      }
      qexpr = "(" + qs + " / swd)";
    } else {
      s += "  {\n";
      qexpr = series ? "0.f" : hexf(b.q);
    }
    std::snprintf(buf, sizeof buf, "    c->bc_q[%d] = %s; c->bc_z[%d] = s_zv[%d];\n  }\n", i,
                  qexpr.c_str(), i, i);
    s += buf;
  }
  s += "  if (mode == 1) {\n"
       "    int req = 0;\n"
       "    if (spos < nsave && dur >= next_save) { c->save_pos = spos + 1; req |= FIP_HALT_DUMP; }\n"
       "    if (!(dur < target)) req |= FIP_HALT_FINISHED;\n"
       "    c->halt_req[par] = req;\n"
       "    if (!(req & FIP_HALT_FINISHED) && vdt) c->dt = fip_control_dt(dt, cflp);\n"
       "  }\n"
       "}\n\n";

  // ---- fip_bc_apply ----
  std::vector<int> stage, lim;
  {
    std::vector<fip_bc> pub(n);
    for (int i = 0; i < n; i++) {
      pub[i] = fip_bc{};
      pub[i].type = bcs[i].type; pub[i].side = bcs[i].side;
      pub[i].from = bcs[i].from; pub[i].to = bcs[i].to;
    }
    fip_internal::bc_stages(pub.data(), n, W, H, stage, lim);
  }
  const int nstages = stage.empty() ? 0 : std::max(0, *std::max_element(stage.begin(), stage.end()) + 1);
  s += "extern \"C\" __global__ __launch_bounds__(256) void fip_bc_apply(fip_grid g, int mode, int par,\n"
       "                                                               int stage, int napply) {\n"
       "  fip_ctl* c = g.ctl;\n"
       "  if ((int)blockIdx.x >= napply) {\n"
       "    const int blk = blockIdx.x - napply, nblk = gridDim.x - napply;\n"
       "    if (mode == 0) fip_schedule(g, par, blk, nblk);\n"
       "    else if (mode == 1 && !c->halt_req[par ^ 1]) fip_schedule(g, par ^ 1, blk, nblk);\n"
       "    return;\n"
       "  }\n"
       "  if (mode == 1 && c->halt_req[par ^ 1]) return;\n"
       "  const int cur = mode == 1 ? (par ^ 1) : par;\n"
       "  float* h = g.h[cur];\n  float* qx = g.qx[cur];\n  float* qy = g.qy[cur];\n"
       "  const float* sohle = g.bed;\n"
       "  const int item = blockIdx.x * blockDim.x + threadIdx.x;\n"
       "  int which = -1, tid = 0, side = 0;\n  (void)side;\n"
       "  switch (stage) {\n";
  if (stage_items) stage_items->assign(nstages, 0);
  for (int st = 0; st < nstages; st++) {
    int total = 0;
    for (int i = 0; i < n; i++)
      if (stage[i] == st && lim[i] >= bcs[i].from) total += lim[i] - bcs[i].from + 1;
    std::snprintf(buf, sizeof buf, "  case %d:  // stage %d: %d border positions\n", st, st, total);
    s += buf;
    total = 0;
    bool first = true;
    for (int i = 0; i < n; i++) {
      if (stage[i] != st || lim[i] < bcs[i].from) continue;
      const int cnt = lim[i] - bcs[i].from + 1;
      std::snprintf(buf, sizeof buf,
                    "    %sif (item < %d) { which = %d; tid = %d + (item - %d); side = %d; }\n",
                    first ? "" : "else ", total + cnt, i, bcs[i].from, total, bcs[i].side);
      s += buf;
      first = false;
      total += cnt;
    }
    if (stage_items) (*stage_items)[st] = total;
    s += "    break;\n";
  }
  s += "  default: break;\n  }\n"
       "  if (which < 0) return;\n"
       "  const float qval = c->bc_q[which], zval = c->bc_z[which];\n"
       "  switch (which) {\n";
  for (int i = 0; i < n; i++) {
    if (lim[i] < bcs[i].from) continue;
    std::snprintf(buf, sizeof buf,
                  "  case %d: fip_bc_cell(%d, %d, %d, qval, zval, h, qx, qy, sohle, g.P, g.W, g.H, tid); break;\n",
                  i, bcs[i].type, bcs[i].side, staggered ? 1 : 0);
    s += buf;
  }
  s += "  default: break;\n  }\n}\n";
  return s;
}

int compile_bc(fip_sim* s, const std::string& src) {
  hiprtcProgram prog;
  std::vector<const char*> hdr_src, hdr_names;
  for (int i = 0; i < kEmbedCount; i++) {
    hdr_src.push_back(kEmbedSrc[i]);
    hdr_names.push_back(kEmbedNames[i]);
  }
  if (hiprtcCreateProgram(&prog, src.c_str(), "fip_bc_kernel.hip", kEmbedCount, hdr_src.data(),
                          hdr_names.data()) != HIPRTC_SUCCESS)
    return fail(FIP_E_RTC, "hiprtcCreateProgram failed");
  std::string archopt = "--offload-arch=" + s->arch;
  const char* opts[] = {archopt.c_str(), "-O3", "-std=c++17", "-ffp-contract=off",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    return fail(FIP_E_RTC, "hipRTC compile of the boundary kernel failed:\n" + log);
  }
  size_t code_size = 0;
  hiprtcGetCodeSize(prog, &code_size);
  std::vector<char> code(code_size);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  if (s->module) {
    hipModuleUnload(s->module);
    s->module = nullptr;
  }
  HIPTRY(hipModuleLoadData(&s->module, code.data()));
  HIPTRY(hipModuleGetFunction(&s->bc_post_fn, s->module, "fip_bc_post"));
  HIPTRY(hipModuleGetFunction(&s->bc_apply_fn, s->module, "fip_bc_apply"));
  return FIP_OK;
}

void destroy_graphs(fip_sim* s) {
  for (auto& c : s->graphs) hipGraphExecDestroy(c.exec);
  s->graphs.clear();
}

int one_step(fip_sim* s, int par, int flood = 0) {
  int rc = launch_stencil(s, par, flood);
  if (rc) return rc;
  rc = launch_bc(s, BC_POST, par);
  if (rc) return rc;
  return s->post_hook ? s->post_hook(s->stream, par) : FIP_OK;
}

// Captured chunks kept per simulation: a run uses at most three lengths (the steps up to
// the next multiple of kStepsPerGraph, whole chunks, the rest), each at one or two parities
// and with or without the flood plains in its first step.
constexpr size_t kMaxGraphs = 8;

// Flood plains (src/main.cpp:125-128: after the step that reaches iteration % 100 == 0).
// In dense mode the next stencil step, which reads that state, takes them for the cells it
// owns (flood_first); in masked mode, and when a run returns at such an iteration, the
// separate pass does (it is idempotent: max(hmax, h) of the same state).
bool fused_flood(const fip_sim* s) { return !s->masked; }

// The graph of `steps` steps (1..kStepsPerGraph) starting at parity par; flood_first: the
// first step takes the flood plains. In masked mode it ends with the flood-plain pass,
// which tests iteration % 100 itself. Captured and instantiated on first use; nullptr
// when capture failed.
hipGraphExec_t chunk_graph(fip_sim* s, int par, int steps, bool flood_first) {
  for (auto& c : s->graphs)
    if (c.par == par && c.steps == steps && c.flood_first == flood_first) return c.exec;
  hipGraph_t graph;
  if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
  int rc = FIP_OK;
  for (int k = 0; k < steps && rc == FIP_OK; k++) rc = one_step(s, par ^ (k & 1), k == 0 && flood_first);
  if (rc == FIP_OK && !fused_flood(s)) rc = launch_misc(s, 3, 1, 0);
  const hipError_t e = hipStreamEndCapture(s->stream, &graph);
  if (rc != FIP_OK || e != hipSuccess) return nullptr;
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ei != hipSuccess) return nullptr;
  // upload now, so that the chunk's first launch does not pay for it
  if (hipGraphUpload(exec, s->stream) != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess) {
    hipGraphExecDestroy(exec);
    return nullptr;
  }
  if (s->graphs.size() == kMaxGraphs) {
    hipGraphExecDestroy(s->graphs.front().exec);
    s->graphs.erase(s->graphs.begin());
  }
  s->graphs.push_back({par, steps, flood_first, exec});
  return exec;
}

// Whether the chunk that starts at iteration it0 takes the flood plains in its first step.
bool flood_first_at(const fip_sim* s, long long it0) {
  return fused_flood(s) && it0 > 0 && it0 % kStepsPerGraph == 0 && s->flood_done_iter != it0;
}

// The chunks fip_simulate cuts `nsteps` steps from iteration it0 into: up to the next
// multiple of kStepsPerGraph (the flood-plain cadence), then whole chunks. Calls
// f(parity, steps, flood_first) for each distinct chunk in order.
template <class F>
void for_each_chunk(const fip_sim* s, int it0, long long nsteps, F f) {
  long long it = it0, left = nsteps;
  int seen = 0;
  int last_par = -1, last_k = -1, last_fl = -1;
  while (left > 0 && seen < 8) {
    const long long k = std::min<long long>(kStepsPerGraph - it % kStepsPerGraph, left);
    const int par = (int)(it & 1);
    const int fl = flood_first_at(s, it) ? 1 : 0;
    if (par != last_par || k != last_k || fl != last_fl) {
      f(par, (int)k, fl != 0);
      seen++;
      last_par = par;
      last_k = (int)k;
      last_fl = fl;
    }
    // whole chunks keep their parity (kStepsPerGraph is even): skip to the last one
    if (k == kStepsPerGraph && left >= 2 * kStepsPerGraph && it >= kStepsPerGraph) {
      const long long whole = left / kStepsPerGraph;
      it += (whole - 1) * kStepsPerGraph;
      left -= (whole - 1) * kStepsPerGraph;
      continue;
    }
    it += k;
    left -= k;
  }
}

int validate_bc(const fip_desc& d, const fip_bc& b) {
  if (b.type < 0 || b.type > 3) return fail(FIP_E_INVALID, "invalid boundary condition type");
  if (b.side < 0 || b.side > 3) return fail(FIP_E_INVALID, "invalid boundary condition side");
  const int ext = (b.side == FIP_SIDE_LEFT || b.side == FIP_SIDE_RIGHT) ? d.H : d.W;
  if (b.from < 0 || b.to >= ext)
    return fail(FIP_E_INVALID, "boundary condition range [" + std::to_string(b.from) + ", " +
                                   std::to_string(b.to) + "] outside the side's " +
                                   std::to_string(ext) + " cells");
  for (int k = 1; k < b.zn; k++)
    if (!(b.zt[k] > b.zt[k - 1])) return fail(FIP_E_INVALID, "z series times must increase");
  for (int k = 1; k < b.qn; k++)
    if (!(b.qt[k] > b.qt[k - 1])) return fail(FIP_E_INVALID, "q series times must increase");
  return FIP_OK;
}

}  // namespace

extern "C" {

const char* fip_last_error(void) { return g_err.c_str(); }

int fip_create(const fip_desc* d, fip_sim** out) {
  if (!d || !out) return fail(FIP_E_INVALID, "null argument");
  *out = nullptr;
  if (d->W < 4 || d->H < 4) return fail(FIP_E_INVALID, "W and H must be at least 4");
  if (!d->bed) return fail(FIP_E_INVALID, "bed (terrain) is required");
  if (!(d->dx > 0.f)) return fail(FIP_E_INVALID, "dx must be positive");
  if (!(d->dt > 0.f)) return fail(FIP_E_INVALID, "dt must be positive");
  const int R = d->tile_rows == 0 ? 64 : d->tile_rows;
  if (R != 16 && R != 32 && R != 64) return fail(FIP_E_INVALID, "tile_rows must be 16, 32 or 64");
  if (d->mode != FIP_MODE_DENSE && d->mode != FIP_MODE_MASKED)
    return fail(FIP_E_INVALID, "invalid mode");
  if (d->variant != FIP_EXACT && d->variant != FIP_FAST) return fail(FIP_E_INVALID, "invalid variant");

  fip_sim* s = new fip_sim();
  s->d = *d;
  s->d.tile_rows = R;
  s->dev = d->device;
  s->masked = d->mode == FIP_MODE_MASKED;
  s->fast = d->variant == FIP_FAST;
  s->wpb = default_wpb(s);
  auto cleanup = [&](int rc) {
    fip_destroy(s);
    return rc;
  };
  if (hipSetDevice(s->dev) != hipSuccess) return cleanup(fail(FIP_E_HIP, "hipSetDevice failed (no GPU?)"));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, s->dev) != hipSuccess)
    return cleanup(fail(FIP_E_HIP, "hipGetDeviceProperties failed"));
  s->arch = prop.gcnArchName;
  s->graph_failed = !fip_internal::graphs_enabled();
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(FIP_E_HIP, "hipStreamCreate failed"));
  if (hipHostMalloc((void**)&s->pin, 16 * sizeof(int), hipHostMallocDefault) != hipSuccess)
    return cleanup(fail(FIP_E_HIP, "hipHostMalloc failed"));

  const int W = d->W, H = d->H;
  const int P = ((kMarginCols + W + kMarginCols) + 63) / 64 * 64;
  const int rows = H + 2 * kMarginRows;
  s->field_elems = (size_t)P * rows;
  s->interior_off = (long)kMarginRows * P + kMarginCols;
  if (s->field_elems * sizeof(float) > (size_t)0x7fffffff)
    return cleanup(fail(FIP_E_INVALID, "grid too large: a field must stay below 2 GiB"));
  fip_grid& g = s->g;
  g.W = W;
  g.H = H;
  g.P = P;
  g.margin = (int)s->interior_off;
  g.field_bytes = (int)(s->field_elems * sizeof(float));
  g.tile_rows = R;
  g.tiles_x = (W + kTileCols - 1) / kTileCols;
  g.tiles_y = (H + R - 1) / R;
  s->ntiles = g.tiles_x * g.tiles_y;
  g.dx = d->dx;
  g.no_data = d->no_data;
  g.kst_uniform = d->kst;
  const int nfield = d->kst_var ? 9 : 8;
  for (int k = 0; k < nfield; k++) {
    if (hipMalloc(&s->base_alloc[k], s->field_elems * sizeof(float)) != hipSuccess)
      return cleanup(fail(FIP_E_HIP, "hipMalloc of a field failed"));
    if (hipMemset(s->base_alloc[k], 0, s->field_elems * sizeof(float)) != hipSuccess)
      return cleanup(fail(FIP_E_HIP, "hipMemset failed"));
  }
  float* in[9];
  for (int k = 0; k < nfield; k++) in[k] = s->base_alloc[k] + s->interior_off;
  g.h[0] = in[0]; g.h[1] = in[1];
  g.qx[0] = in[2]; g.qx[1] = in[3];
  g.qy[0] = in[4]; g.qy[1] = in[5];
  g.hmax = in[6];
  g.bed = in[7];
  g.kst = d->kst_var ? in[8] : nullptr;
  // initial state into both ping-pong sides (src/simdata.cpp:82-90), hmax starts as h
  const size_t pitch = (size_t)P * sizeof(float), wb = (size_t)W * sizeof(float);
  std::vector<float> zeros;
  auto up = [&](float* dst, const float* src) -> int {
    if (!src) return FIP_OK;
    if (hipMemcpy2D(dst, pitch, src, wb, wb, H, hipMemcpyHostToDevice) != hipSuccess)
      return fail(FIP_E_HIP, "hipMemcpy2D upload failed");
    return FIP_OK;
  };
  int rc = FIP_OK;
  rc |= up(g.h[0], d->h); rc |= up(g.h[1], d->h); rc |= up(g.hmax, d->h);
  rc |= up(g.qx[0], d->qx); rc |= up(g.qx[1], d->qx);
  rc |= up(g.qy[0], d->qy); rc |= up(g.qy[1], d->qy);
  rc |= up((float*)g.bed, d->bed);
  if (d->kst_var) rc |= up((float*)g.kst, d->kst_var);
  if (rc) return cleanup(FIP_E_HIP);
  if (d->kst_var) {
    // kSt by class: the stencil then reads one byte per cell instead of four and looks
    // the value up, which gives the same float
    std::unordered_map<uint32_t, int> cls;
    std::vector<float> tab;
    std::vector<unsigned char> idx((size_t)W * H);
    bool ok = true;
    for (size_t i = 0; ok && i < (size_t)W * H; i++) {
      uint32_t bits;
      std::memcpy(&bits, &d->kst_var[i], 4);
      auto it = cls.find(bits);
      if (it == cls.end()) {
        if ((int)tab.size() == FIP_KST_CLASSES) { ok = false; break; }
        it = cls.emplace(bits, (int)tab.size()).first;
        tab.push_back(d->kst_var[i]);
      }
      idx[i] = (unsigned char)it->second;
    }
    if (ok) {
      while ((int)tab.size() < FIP_KST_CLASSES) tab.push_back(tab[0]);
      if (hipMalloc(&s->kidx_alloc, s->field_elems) != hipSuccess ||
          hipMemset(s->kidx_alloc, 0, s->field_elems) != hipSuccess ||
          hipMalloc(&s->ktab_alloc, sizeof(float) * FIP_KST_CLASSES) != hipSuccess)
        return cleanup(fail(FIP_E_HIP, "kSt class allocation failed"));
      g.kidx = s->kidx_alloc + s->interior_off;
      g.ktab = s->ktab_alloc;
      s->stats.kst_classes = (int)cls.size();
      if (hipMemcpy2D(s->kidx_alloc + s->interior_off, (size_t)P, idx.data(), (size_t)W, (size_t)W, H,
                      hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(s->ktab_alloc, tab.data(), sizeof(float) * FIP_KST_CLASSES, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(FIP_E_HIP, "kSt class upload failed"));
    }
  }
  // masks: every tile wet in every row (src/simdata.cpp:77-79)
  std::vector<unsigned long long> masks(s->ntiles);
  for (int t = 0; t < s->ntiles; t++) {
    const int ty = t / g.tiles_x;
    const int nr = std::min(R, H - ty * R);
    masks[t] = nr >= 64 ? ~0ull : ((1ull << nr) - 1ull);
  }
  if (hipMalloc(&g.mask, sizeof(unsigned long long) * s->ntiles) != hipSuccess ||
      hipMemcpy(g.mask, masks.data(), sizeof(unsigned long long) * s->ntiles, hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(FIP_E_HIP, "mask allocation failed"));
  for (int k = 0; k < 2; k++)
    if (hipMalloc(&g.list[k], sizeof(unsigned) * s->ntiles) != hipSuccess)
      return cleanup(fail(FIP_E_HIP, "tile list allocation failed"));
  fip_ctl* hc = new fip_ctl();
  std::memset(hc, 0, sizeof(fip_ctl));
  hc->dt = d->dt;
  hc->cfl_prev = 1.0f;
  const float one = 1.0f;
  std::memcpy(&hc->cfl_acc[0], &one, 4);
  std::memcpy(&hc->cfl_acc[1], &one, 4);
  hc->variable_dt = d->variable_dt;
  hc->target = 0.0;
  bool ok = hipMalloc(&g.ctl, sizeof(fip_ctl)) == hipSuccess &&
            hipMemcpy(g.ctl, hc, sizeof(fip_ctl), hipMemcpyHostToDevice) == hipSuccess;
  delete hc;
  if (!ok) return cleanup(fail(FIP_E_HIP, "control block allocation failed"));
  const int cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  s->cus = cus;
  s->stencil_blocks = std::max(1, cus * std::max(1, s->bpc));
  s->sched_blocks = std::max(1, std::min((s->ntiles + 255) / 256, cus));
  if (int rc2 = plan_dense(s)) return cleanup(rc2);
  if (hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
    return cleanup(fail(FIP_E_HIP, "hipEventCreate failed"));
  s->stats.tiles_total = s->ntiles;
  if (hipDeviceSynchronize() != hipSuccess) return cleanup(fail(FIP_E_HIP, "device sync failed"));
  *out = s;
  return FIP_OK;
}

int fip_destroy(fip_sim* s) {
  if (!s) return FIP_OK;
  hipSetDevice(s->dev);
  if (s->stream) hipStreamSynchronize(s->stream);
  fip_internal::closure_forget(s, nullptr);
  s->post_hook = nullptr;
  destroy_graphs(s);
  if (s->kidx_alloc) hipFree(s->kidx_alloc);
  if (s->ktab_alloc) hipFree(s->ktab_alloc);
  for (auto& p : s->base_alloc)
    if (p) hipFree(p);
  if (s->g.mask) hipFree(s->g.mask);
  for (auto& l : s->g.list)
    if (l) hipFree(l);
  if (s->g.ctl) hipFree(s->g.ctl);
  if (s->d_series) hipFree(s->d_series);
  if (s->module) hipModuleUnload(s->module);
  if (s->ev0) hipEventDestroy(s->ev0);
  if (s->ev1) hipEventDestroy(s->ev1);
  if (s->stream) hipStreamDestroy(s->stream);
  if (s->pin) hipHostFree(s->pin);
  delete s;
  return FIP_OK;
}

static int collect_bcs(const fip_desc& d, const fip_bc* bcs, int n, std::vector<BCDesc>& v,
                       std::vector<float>& table) {
  if (n > 0 && !bcs) return fail(FIP_E_INVALID, "null argument");
  if (n < 0 || n > FIP_MAX_BC) return fail(FIP_E_INVALID, "too many boundary conditions");
  for (int i = 0; i < n; i++) {
    int rc = validate_bc(d, bcs[i]);
    if (rc) return rc;
    BCDesc b{bcs[i].type, bcs[i].side, bcs[i].from, bcs[i].to, bcs[i].z, bcs[i].q, 0, 0, 0, 0};
    if (bcs[i].zn > 0) {
      b.zoff = (int)table.size();
      b.zn = bcs[i].zn;
      table.insert(table.end(), bcs[i].zt, bcs[i].zt + b.zn);
      table.insert(table.end(), bcs[i].zv, bcs[i].zv + b.zn);
    }
    if (bcs[i].qn > 0) {
      b.qoff = (int)table.size();
      b.qn = bcs[i].qn;
      table.insert(table.end(), bcs[i].qt, bcs[i].qt + b.qn);
      table.insert(table.end(), bcs[i].qv, bcs[i].qv + b.qn);
    }
    v.push_back(b);
  }
  return FIP_OK;
}

int fip_bc_source(const fip_bc* bcs, int n, int W, int H, int staggered, char* buf, size_t cap,
                  size_t* len) {
  fip_desc d{};
  d.W = W;
  d.H = H;
  std::vector<BCDesc> v;
  std::vector<float> table;
  int rc = collect_bcs(d, bcs, n, v, table);
  if (rc) return rc;
  const std::string src = bc_source(v, staggered != 0, W, H, nullptr);
  if (len) *len = src.size();
  if (buf && cap > 0) {
    const size_t k = std::min(cap - 1, src.size());
    std::memcpy(buf, src.data(), k);
    buf[k] = '\0';
  }
  return FIP_OK;
}

int fip_bc_generate(fip_sim* s, const fip_bc* bcs, int n) {
  if (!s) return fail(FIP_E_INVALID, "null argument");
  HIPTRY(hipSetDevice(s->dev));
  std::vector<BCDesc> v;
  std::vector<float> table;
  int rc = collect_bcs(s->d, bcs, n, v, table);
  if (rc) return rc;
  if (s->d_series) {
    hipFree(s->d_series);
    s->d_series = nullptr;
  }
  if (!table.empty()) {
    HIPTRY(hipMalloc(&s->d_series, table.size() * sizeof(float)));
    HIPTRY(hipMemcpy(s->d_series, table.data(), table.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  s->g.series = s->d_series;
  std::vector<int> items;
  rc = compile_bc(s, bc_source(v, s->d.staggered != 0, s->d.W, s->d.H, &items));
  if (!rc) s->bc_stage_items = items;
  if (rc) return rc;
  // reset the sumWaterDepths cache
  int zeros[FIP_MAX_BC] = {0};
  rc = ctl_write(s, CTL_OFF(swd_valid), zeros, FIP_MAX_BC);
  if (rc) return rc;
  s->nbc = n;
  s->bc_ready = true;
  destroy_graphs(s);
  return FIP_OK;
}

int fip_bc_launch(fip_sim* s, const float* params) {
  if (!s) return fail(FIP_E_INVALID, "null argument");
  HIPTRY(hipSetDevice(s->dev));
  if (s->nbc > 0) {
    std::vector<float> q(s->nbc), z(s->nbc);
    for (int i = 0; i < s->nbc; i++) {
      q[i] = params[2 * i];
      z[i] = params[2 * i + 1];
    }
    int rc = ctl_write(s, CTL_OFF(bc_q), q.data(), s->nbc);
    if (!rc) rc = ctl_write(s, CTL_OFF(bc_z), z.data(), s->nbc);
    if (rc) return rc;
  }
  return launch_bc(s, BC_APPLY_HOST, s->host_iter & 1);
}

int fip_bc_apply(fip_sim* s) {
  if (!s) return fail(FIP_E_INVALID, "null argument");
  HIPTRY(hipSetDevice(s->dev));
  return launch_bc(s, BC_APPLY_DEVICE, s->host_iter & 1);
}

int fip_bc_params(fip_sim* s, float* params, int n) {
  if (!s || !params) return fail(FIP_E_INVALID, "null argument");
  n = std::min(n, s->nbc);
  std::vector<float> q(n), z(n);
  int rc = ctl_read(s, CTL_OFF(bc_q), q.data(), n);
  if (!rc) rc = ctl_read(s, CTL_OFF(bc_z), z.data(), n);
  for (int i = 0; i < n; i++) {
    params[2 * i] = q[i];
    params[2 * i + 1] = z[i];
  }
  return rc;
}

int fip_control_timestep(fip_sim* s) {
  if (!s) return fail(FIP_E_INVALID, "null argument");
  HIPTRY(hipSetDevice(s->dev));
  return launch_misc(s, 0, 0, 1);
}

int fip_compute_step(fip_sim* s) {
  if (!s) return fail(FIP_E_INVALID, "null argument");
  HIPTRY(hipSetDevice(s->dev));
  const int par = s->host_iter & 1;
  if (s->masked && !s->started) {
    // first step of a fine-grained run: build the initial list from the all-wet masks
    unsigned zero = 0;
    int rc = ctl_write(s, CTL_OFF(list_count[0]), &zero);
    if (!rc) rc = ctl_write(s, CTL_OFF(list_count[1]), &zero);
    if (!rc) rc = launch_misc(s, 2, par, s->sched_blocks);
    if (rc) return rc;
    s->started = true;
  }
  int rc = launch_stencil(s, par);
  if (!rc) rc = launch_misc(s, 1, par, 1);
  if (!rc && s->masked) rc = launch_misc(s, 2, par ^ 1, s->sched_blocks);
  if (rc) return rc;
  s->host_iter++;
  s->stats.steps++;
  return FIP_OK;
}

int fip_floodplain(fip_sim* s) {
  if (!s) return fail(FIP_E_INVALID, "null argument");
  HIPTRY(hipSetDevice(s->dev));
  return launch_misc(s, 3, 0, 0);
}

int fip_set_run(fip_sim* s, double target, int sampling_interval, const int* sx, const int* sy,
                const int* st, int npoints, const double* save_times, int nsave) {
  if (!s) return fail(FIP_E_INVALID, "null argument");
  if (npoints < 0 || npoints > FIP_MAX_SAMPLE_POINTS)
    return fail(FIP_E_INVALID, "too many sampling points (max " + std::to_string(FIP_MAX_SAMPLE_POINTS) + ")");
  if (nsave < 0 || nsave > FIP_MAX_SAVE)
    return fail(FIP_E_INVALID, "too many save_state times (max " + std::to_string(FIP_MAX_SAVE) + ")");
  for (int p = 0; p < npoints; p++)
    if (sx[p] < 0 || sx[p] >= s->d.W || sy[p] < 0 || sy[p] >= s->d.H || st[p] < 0 || st[p] > 3)
      return fail(FIP_E_INVALID, "sampling point outside the grid");
  HIPTRY(hipSetDevice(s->dev));
  fip_ctl* hc = new fip_ctl();
  int rc = ctl_read(s, 0, hc);
  if (rc) {
    delete hc;
    return rc;
  }
  hc->target = target;
  hc->sampling_interval = sampling_interval;
  hc->sample_timer = 0.0;  // Sampling::timer{duration} (src/simdata.cpp:41)
  hc->pre_sampled = 0;
  hc->nsample_points = npoints;
  for (int p = 0; p < npoints; p++) {
    hc->sp_x[p] = sx[p];
    hc->sp_y[p] = sy[p];
    hc->sp_type[p] = st[p];
  }
  std::vector<double> sv(save_times, save_times + nsave);
  std::sort(sv.begin(), sv.end());
  for (int k = 0; k < nsave; k++) hc->save_times[k] = sv[k];
  hc->nsave = nsave;
  hc->save_pos = 0;
  rc = ctl_write(s, 0, hc);
  delete hc;
  s->npoints = npoints;
  s->run_set = true;
  return rc;
}

int fip_simulate(fip_sim* s, long long nsteps, int* event, long long* steps_done) {
  if (!s || !event) return fail(FIP_E_INVALID, "null argument");
  *event = FIP_EV_NONE;
  if (steps_done) *steps_done = 0;
  if (!s->run_set) return fail(FIP_E_STATE, "call fip_set_run first");
  HIPTRY(hipSetDevice(s->dev));
  int rc;
  // pending stop requests: a finished run stays finished; a handled dump is cleared
  CtlStatus c0;
  if ((rc = ctl_status(s, &c0))) return rc;
  const int* req = c0.halt_req;
  if ((req[0] | req[1]) & FIP_HALT_FINISHED) {
    *event = FIP_EV_FINISHED;
    return FIP_OK;
  }
  if (req[0] | req[1]) {
    int z[2] = {0, 0};
    if ((rc = ctl_write(s, CTL_OFF(halt_req), z, 2))) return rc;
  }
  if (!s->started) {
    double t, target;
    if ((rc = ctl_read(s, CTL_OFF(duration), &t))) return rc;
    if ((rc = ctl_read(s, CTL_OFF(target), &target))) return rc;
    if (!(t < target)) {
      *event = FIP_EV_FINISHED;
      return FIP_OK;
    }
    if ((rc = launch_bc(s, BC_PRE0, s->host_iter & 1))) return rc;
    s->started = true;
  }
  long long done = 0;
  while (done < nsteps) {
    const int it0 = s->host_iter;
    const long long to_flood = kStepsPerGraph - (it0 % kStepsPerGraph);
    const long long k = std::min<long long>(to_flood, nsteps - done);
    const bool flood_first = flood_first_at(s, it0);
    hipGraphExec_t gx = nullptr;
    if (!s->timing && !s->graph_failed) {
      gx = chunk_graph(s, it0 & 1, (int)k, flood_first);
      if (!gx) {
        s->graph_failed = true;
        hipGetLastError();
      }
    }
    if (gx) {
      HIPTRY(hipGraphLaunch(gx, s->stream));
      s->stats.stencil_launches += k;
    } else {
      for (long long i = 0; i < k; i++)
        if ((rc = one_step(s, (it0 + (int)i) & 1, i == 0 && flood_first))) return rc;
      if (!fused_flood(s) && (rc = launch_misc(s, 3, 1, 0))) return rc;
    }
    CtlStatus cs;
    if ((rc = ctl_status(s, &cs))) return rc;
    const int it1 = cs.iteration;
    s->stats.steps += it1 - it0;
    done += it1 - it0;
    s->host_iter = it1;
    const int r = cs.halt_req[0] | cs.halt_req[1];
    if (r & FIP_HALT_DUMP) {
      *event = FIP_EV_DUMP;
      break;
    }
    if (r & FIP_HALT_FINISHED) {
      *event = FIP_EV_FINISHED;
      break;
    }
    if (s->npoints > 0 && cs.nsample_rows > FIP_SAMPLE_LOG_ROWS - kStepsPerGraph - 1) {
      *event = FIP_EV_SAMPLES;
      break;
    }
  }
  if (steps_done) *steps_done = done;
  // the run returns on a flood-plain iteration: take them now (the next step would)
  if (fused_flood(s) && s->host_iter > 0 && s->host_iter % kStepsPerGraph == 0 &&
      s->flood_done_iter != s->host_iter) {
    if ((rc = launch_misc(s, 3, 1, 0))) return rc;
    s->flood_done_iter = s->host_iter;
  }
  if (s->masked) {
    unsigned cnt[2];
    if ((rc = ctl_read(s, CTL_OFF(list_count), cnt, 2))) return rc;
    s->stats.tiles_computed = cnt[s->host_iter & 1];
  } else {
    s->stats.tiles_computed = s->ntiles;
  }
  return FIP_OK;
}

int fip_prepare(fip_sim* s, long long nsteps) {
  if (!s) return fail(FIP_E_INVALID, "null argument");
  if (!s->run_set) return fail(FIP_E_STATE, "call fip_set_run first");
  if (!s->bc_ready) return fail(FIP_E_STATE, "boundary kernel not generated (call fip_bc_generate)");
  if (nsteps < 0) return fail(FIP_E_INVALID, "nsteps must not be negative");
  HIPTRY(hipSetDevice(s->dev));
  if (s->timing || s->graph_failed) return FIP_OK;
  bool ok = true;
  for_each_chunk(s, s->host_iter, nsteps, [&](int par, int k, bool fl) {
    if (ok && !chunk_graph(s, par, k, fl)) ok = false;
  });
  if (!ok) {
    s->graph_failed = true;
    hipGetLastError();
  }
  return FIP_OK;
}

int fip_read_samples(fip_sim* s, double* t, float* vals, int max_rows, int* nrows) {
  if (!s || !nrows) return fail(FIP_E_INVALID, "null argument");
  int n = 0, ovf = 0;
  int rc = ctl_read(s, CTL_OFF(nsample_rows), &n);
  if (!rc) rc = ctl_read(s, CTL_OFF(sample_overflow), &ovf);
  if (rc) return rc;
  if (ovf) return fail(FIP_E_STATE, "sample log overflow");
  n = std::min(n, max_rows);
  if (n > 0) {
    std::vector<double> tt(n);
    std::vector<float> vv((size_t)n * FIP_MAX_SAMPLE_POINTS);
    if ((rc = ctl_read(s, CTL_OFF(sample_t), tt.data(), n))) return rc;
    if ((rc = ctl_read(s, CTL_OFF(sample_v), vv.data(), vv.size()))) return rc;
    for (int r = 0; r < n; r++) {
      t[r] = tt[r];
      for (int p = 0; p < s->npoints; p++)
        vals[(size_t)r * s->npoints + p] = vv[(size_t)r * FIP_MAX_SAMPLE_POINTS + p];
    }
    // drain: move rows beyond n (none when max_rows is large enough) and reset
    int zero = 0;
    if ((rc = ctl_write(s, CTL_OFF(nsample_rows), &zero))) return rc;
  }
  *nrows = n;
  return FIP_OK;
}

int fip_get_time(fip_sim* s, double* duration, float* dt, int* iteration) {
  if (!s) return fail(FIP_E_INVALID, "null argument");
  HIPTRY(hipSetDevice(s->dev));
  int rc;
  if (duration && (rc = ctl_read(s, CTL_OFF(duration), duration))) return rc;
  if (dt && (rc = ctl_read(s, CTL_OFF(dt), dt))) return rc;
  if (iteration && (rc = ctl_read(s, CTL_OFF(iteration), iteration))) return rc;
  return FIP_OK;
}

int fip_set_dt(fip_sim* s, float dt) {
  if (!s) return fail(FIP_E_INVALID, "null argument");
  HIPTRY(hipSetDevice(s->dev));
  return ctl_write(s, CTL_OFF(dt), &dt);
}

static float* field_ptr(fip_sim* s, int field, int side) {
  const int k = (s->host_iter & 1) ^ (side ? 1 : 0);
  switch (field) {
    case FIP_FIELD_H: return s->g.h[k];
    case FIP_FIELD_QX: return s->g.qx[k];
    case FIP_FIELD_QY: return s->g.qy[k];
    case FIP_FIELD_HMAX: return s->g.hmax;
    case FIP_FIELD_BED: return (float*)s->g.bed;
    case FIP_FIELD_KST: return (float*)s->g.kst;
    default: return nullptr;
  }
}

int fip_download(fip_sim* s, int field, int side, float* dst) {
  if (!s || !dst) return fail(FIP_E_INVALID, "null argument");
  HIPTRY(hipSetDevice(s->dev));
  HIPTRY(hipStreamSynchronize(s->stream));
  float* src = field_ptr(s, field, side);
  if (!src) return fail(FIP_E_INVALID, "invalid field");
  const size_t wb = (size_t)s->d.W * sizeof(float);
  HIPTRY(hipMemcpy2D(dst, wb, src, (size_t)s->g.P * sizeof(float), wb, s->d.H, hipMemcpyDeviceToHost));
  return FIP_OK;
}

int fip_upload(fip_sim* s, int field, int side, const float* src) {
  if (!s || !src) return fail(FIP_E_INVALID, "null argument");
  HIPTRY(hipSetDevice(s->dev));
  HIPTRY(hipStreamSynchronize(s->stream));
  float* dst = field_ptr(s, field, side);
  if (!dst) return fail(FIP_E_INVALID, "invalid field");
  const size_t wb = (size_t)s->d.W * sizeof(float);
  HIPTRY(hipMemcpy2D(dst, (size_t)s->g.P * sizeof(float), src, wb, wb, s->d.H, hipMemcpyHostToDevice));
  return FIP_OK;
}

int fip_tiles(fip_sim* s, int* tx, int* ty, int* tc, int* tr) {
  if (!s) return fail(FIP_E_INVALID, "null argument");
  if (tx) *tx = s->g.tiles_x;
  if (ty) *ty = s->g.tiles_y;
  if (tc) *tc = kTileCols;
  if (tr) *tr = s->g.tile_rows;
  return FIP_OK;
}

int fip_get_masks(fip_sim* s, uint64_t* dst, int n) {
  if (!s || !dst) return fail(FIP_E_INVALID, "null argument");
  HIPTRY(hipSetDevice(s->dev));
  HIPTRY(hipStreamSynchronize(s->stream));
  n = std::min(n, s->ntiles);
  HIPTRY(hipMemcpy(dst, s->g.mask, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
  return FIP_OK;
}

int fip_get_cfl(fip_sim* s, float* cfl) {
  if (!s || !cfl) return fail(FIP_E_INVALID, "null argument");
  return ctl_read(s, CTL_OFF(cfl_prev), cfl);
}

int fip_set_launch(fip_sim* s, int waves_per_block, int blocks_per_cu) {
  if (!s) return fail(FIP_E_INVALID, "null argument");
  if (waves_per_block < 0 || waves_per_block > 16 || blocks_per_cu > 64)
    return fail(FIP_E_INVALID, "waves_per_block must be 0..16 and blocks_per_cu at most 64");
  s->wpb = waves_per_block != 0 ? waves_per_block : default_wpb(s);
  s->bpc = blocks_per_cu < 0 ? 0 : blocks_per_cu;
  s->stencil_blocks = std::max(1, s->cus * std::max(1, s->bpc));
  destroy_graphs(s);
  HIPTRY(hipSetDevice(s->dev));
  return plan_dense(s);
}

int fip_tune(const fip_desc* d, const fip_bc* bcs, int nbc, int steps, fip_tune_result* out) {
  if (!d || !out || (nbc > 0 && !bcs)) return fail(FIP_E_INVALID, "null argument");
  if (steps < 1) return fail(FIP_E_INVALID, "steps must be at least 1");
  static const int kRows[] = {16, 32, 64};
  static const int kWaves[] = {1, 2, 4, 8, 16};
  static const int kBlocksPerCu[] = {0, 4, 8, 16};
  const bool masked = d->mode == FIP_MODE_MASKED;
  fip_tune_result best{};
  best.stencil_ms = 1e30;
  for (int R : kRows) {
    // dense mode marches spans sized at launch: the tile rows do not reach its kernel
    if (!masked && R != (d->tile_rows == 0 ? 64 : d->tile_rows)) continue;
    fip_desc dd = *d;
    dd.tile_rows = R;
    fip_sim* s = nullptr;
    int rc = fip_create(&dd, &s);
    if (!rc) rc = fip_bc_generate(s, bcs, nbc);
    if (!rc) rc = fip_set_run(s, 1e300, 0, nullptr, nullptr, nullptr, 0, nullptr, 0);
    for (int w : kWaves) {
      for (int bpc : kBlocksPerCu) {
        if (rc) break;
        if (!masked && bpc != kBlocksPerCu[0]) continue;
        int ev = 0;
        long long done = 0;
        rc = fip_set_launch(s, w, masked ? bpc : -1);
        if (!rc) rc = fip_set_timing(s, 0);
        if (!rc) rc = fip_simulate(s, steps, &ev, &done);
        const fip_stats before = s ? s->stats : fip_stats{};
        if (!rc) rc = fip_set_timing(s, 1);
        if (!rc) rc = fip_simulate(s, steps, &ev, &done);
        if (rc) break;
        const long long n = s->stats.timed_launches - before.timed_launches;
        const double ms = n > 0 ? (s->stats.stencil_ms - before.stencil_ms) / (double)n : 1e30;
        best.tried++;
        if (ms < best.stencil_ms) {
          best.stencil_ms = ms;
          best.tile_rows = R;
          best.waves_per_block = w;
          best.blocks_per_cu = masked ? bpc : -1;
        }
      }
    }
    if (s) fip_destroy(s);
    if (rc) return rc;
  }
  *out = best;
  return FIP_OK;
}

int fip_set_timing(fip_sim* s, int enable) {
  if (!s) return fail(FIP_E_INVALID, "null argument");
  s->timing = enable != 0;
  return FIP_OK;
}

int fip_get_stats(fip_sim* s, fip_stats* out) {
  if (!s || !out) return fail(FIP_E_INVALID, "null argument");
  *out = s->stats;
  out->dense_span = s->masked ? 0 : s->dense_span;
  return FIP_OK;
}

int fip_sync(fip_sim* s) {
  if (!s) return fail(FIP_E_INVALID, "null argument");
  HIPTRY(hipSetDevice(s->dev));
  HIPTRY(hipStreamSynchronize(s->stream));
  return FIP_OK;
}

int fip_check_rcp(int device, float lo, float hi, unsigned long long* nbad, float* first_bad) {
  if (!nbad || !first_bad) return fail(FIP_E_INVALID, "null argument");
  if (!(lo > 0.0f) || !(hi >= lo) || !(hi <= 0x1p64f))
    return fail(FIP_E_INVALID, "the range must satisfy 0 < lo <= hi <= 2^64");
  HIPTRY(hipSetDevice(device));
  unsigned long long* dn = nullptr;
  unsigned* df = nullptr;
  HIPTRY(hipMalloc(&dn, sizeof(*dn)));
  hipError_t e = hipMalloc(&df, sizeof(*df));
  unsigned a, bh;
  std::memcpy(&a, &lo, 4);
  std::memcpy(&bh, &hi, 4);
  const unsigned n = bh - a + 1u;
  const unsigned none = 0xffffffffu;
  if (e == hipSuccess) e = hipMemset(dn, 0, sizeof(*dn));
  if (e == hipSuccess) e = hipMemcpy(df, &none, sizeof(none), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = fip_launch_check_rcp(a, n, dn, df, nullptr);
  unsigned f = none;
  if (e == hipSuccess) e = hipMemcpy(nbad, dn, sizeof(*dn), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&f, df, sizeof(f), hipMemcpyDeviceToHost);
  (void)hipFree(dn);
  if (df) (void)hipFree(df);
  if (e != hipSuccess) return fail(FIP_E_HIP, std::string("fip_check_rcp: ") + hipGetErrorString(e));
  float fb = 0.0f;
  if (f != none) std::memcpy(&fb, &f, 4);
  *first_bad = fb;
  return FIP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// library-internal helpers (fip_internal.h)

namespace fip_internal {

int fail(int code, const std::string& msg) { return ::fail(code, msg); }
const fip_grid& grid(const fip_sim* s) { return s->g; }
int staggered(const fip_sim* s) { return s->d.staggered; }
int device(const fip_sim* s) { return s->dev; }

bool graphs_enabled() {
  const char* v = std::getenv("FIP_GRAPHS");
  return !(v && v[0] == '0' && v[1] == 0);
}

void set_post_hook(fip_sim* s, StepHook hook) {
  s->post_hook = std::move(hook);
  destroy_graphs(s);
}
hipStream_t stream(const fip_sim* s) { return s->stream; }
int cur_side(const fip_sim* s) { return s->host_iter & 1; }
fip_grid& grid_mut(fip_sim* s) { return s->g; }

int refresh(fip_sim* s) {
  int it = 0;
  int rc = ctl_read(s, CTL_OFF(iteration), &it);
  if (rc) return rc;
  s->host_iter = it;
  s->started = true;
  return FIP_OK;
}

int bc_stages(const fip_bc* bcs, int n, int W, int H, std::vector<int>& stage,
              std::vector<int>& lim) {
  // the reference launches one thread per position of the longer side (rounded up to
  // 64), so positions beyond from + threads - 1 are never set (src/bc_kernel_gen.cpp:143-149)
  const int ntid = (std::max(W, H) / 64 + 1) * 64;
  stage.assign(n, -1);
  lim.assign(n, 0);
  for (int i = 0; i < n; i++) lim[i] = std::min(bcs[i].to, bcs[i].from + ntid - 1);
  // the two cells deep strip a BC reads or writes at its positions
  auto strip = [&](int i, int* x0, int* x1, int* y0, int* y1) {
    const fip_bc& b = bcs[i];
    if (b.side == FIP_SIDE_LEFT) { *x0 = 0; *x1 = 1; *y0 = b.from; *y1 = lim[i]; }
    else if (b.side == FIP_SIDE_RIGHT) { *x0 = W - 2; *x1 = W - 1; *y0 = b.from; *y1 = lim[i]; }
    else if (b.side == FIP_SIDE_TOP) { *x0 = b.from; *x1 = lim[i]; *y0 = H - 2; *y1 = H - 1; }
    else { *x0 = b.from; *x1 = lim[i]; *y0 = 0; *y1 = 1; }
  };
  int nstages = 0;
  for (int j = 0; j < n; j++) {
    if (lim[j] < bcs[j].from) continue;
    stage[j] = 0;
    int a0, a1, b0, b1;
    strip(j, &a0, &a1, &b0, &b1);
    for (int i = 0; i < j; i++) {
      if (stage[i] < 0) continue;
      int c0, c1, d0, d1;
      strip(i, &c0, &c1, &d0, &d1);
      if (a0 <= c1 && c0 <= a1 && b0 <= d1 && d0 <= b1) stage[j] = std::max(stage[j], stage[i] + 1);
    }
    nstages = std::max(nstages, stage[j] + 1);
  }
  return nstages;
}

}  // namespace fip_internal
