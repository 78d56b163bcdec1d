// libfiphip: host side of the ensemble C ABI (include/fip_ensemble.h). Each scenario is
// a single-domain simulation (its fields, masks and control block come from
// fip_create); the ensemble adds the global tile list, the flattened BC tables and the
// launches that advance every scenario at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fip_ensemble.h"
#include "fip_ctl.h"
#include "fip_internal.h"

hipError_t fip_launch_ens_stencil_exact(const fip_ens_scn* scn, const unsigned long long* list,
                                        fip_ens_ctl* ec, int par, unsigned cap, int tile_rows,
                                        int kvar, int blocks, int wpb, int persistent,
                                        hipStream_t st);
hipError_t fip_launch_ens_stencil_fast(const fip_ens_scn* scn, const unsigned long long* list,
                                       fip_ens_ctl* ec, int par, unsigned cap, int tile_rows,
                                       int kvar, int blocks, int wpb, int persistent,
                                       hipStream_t st);
hipError_t fip_launch_ens_dense_exact(const fip_ens_scn* scn, const int4* items, int nitems, int kvar, int wpb,
                                      int par, hipStream_t st);
hipError_t fip_launch_ens_dense_fast(const fip_ens_scn* scn, const int4* items, int nitems, int kvar, int wpb,
                                     int par, hipStream_t st);
hipError_t fip_ens_dense_waves_per_cu_exact(int kvar, int wpb, int* waves);
hipError_t fip_ens_dense_waves_per_cu_fast(int kvar, int wpb, int* waves);
hipError_t fip_launch_ens(int which, const fip_ens_scn* scn, int nscn, int total_tiles,
                          const fip_ens_bc* bcs, const float* series, const fip_ens_chunk* chunks,
                          int nchunks, fip_ens_ctl* ec, unsigned long long* list, int mode, int par,
                          int sched_q, int dense, int blocks, hipStream_t st);

using fip_internal::fail;

#define EHIP(expr)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(FIP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
  } while (0)

namespace {

constexpr int kStepsPerGraph = 100;  // the flood-plain cadence

struct ScnBCs {
  std::vector<fip_bc> bcs;          // type/side/from/to (+ fixed z/q)
  std::vector<std::vector<float>> zt, zv, qt, qv;
};

}  // namespace

struct fip_ens {
  int n = 0;
  int dev = 0;
  int R = 32;
  bool fast = false;
  bool dense = false;
  bool kvar = false;
  std::vector<fip_sim*> sims;
  std::vector<ScnBCs> bcl;
  hipStream_t stream = nullptr;
  fip_ens_scn* d_scn = nullptr;
  fip_ens_ctl* d_ctl = nullptr;
  unsigned long long* d_list = nullptr;  // [2][total_tiles]
  int4* d_items = nullptr;  // dense mode, per wavefront: (scenario, first row, last row, tile column)
  int nitems = 0;
  int items_wpb = 0;        // the wpb the items were planned for
  int dense_span = 0;
  std::vector<int> scn_tiles;  // tiles per scenario
  fip_ens_bc* d_bcs = nullptr;
  float* d_series = nullptr;
  std::vector<fip_ens_chunk*> d_chunks;  // per stage
  std::vector<int> nchunks;
  int total_tiles = 0;
  int cus = 256;
  int wpb = 4;  // set at creation: 16 for the dense fast variant, 4 otherwise
  int bpc = 0;  // persistent work-groups per CU (0 = one wavefront per list entry)
  int stencil_blocks = 0;
  int sched_blocks = 0;
  long long iter = 0;
  bool built = false;
  bool started = false;
  bool run_set = false;
  // captured chunks: `steps` steps from parity `par`, then the flood-plain kernel (which
  // tests iteration % 100 itself); most recently built last
  struct Chunk {
    int par, steps;
    hipGraphExec_t exec;
  };
  std::vector<Chunk> graphs;
  bool graph_failed = false;
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  fip_stats stats{};
  fip_internal::StepHook post_hook;  // extra work after each ensemble step
};

namespace {

void destroy_graphs(fip_ens* e) {
  for (auto& c : e->graphs) hipGraphExecDestroy(c.exec);
  e->graphs.clear();
}

void free_tables(fip_ens* e) {
  if (e->d_bcs) hipFree(e->d_bcs);
  if (e->d_series) hipFree(e->d_series);
  for (auto* p : e->d_chunks)
    if (p) hipFree(p);
  e->d_bcs = nullptr;
  e->d_series = nullptr;
  e->d_chunks.clear();
  e->nchunks.clear();
}

// Flattens the BC lists of all scenarios into device tables and the per-stage chunk
// lists, and uploads the scenario descriptors.
int build(fip_ens* e) {
  destroy_graphs(e);
  free_tables(e);
  std::vector<fip_ens_bc> bcs;
  std::vector<float> series;
  std::vector<fip_ens_scn> scn(e->n);
  std::vector<std::vector<fip_ens_chunk>> chunks;
  int tile_base = 0;
  e->scn_tiles.assign(e->n, 0);
  for (int s = 0; s < e->n; s++) {
    const fip_grid& g = fip_internal::grid(e->sims[s]);
    scn[s].g = g;
    scn[s].bc0 = (int)bcs.size();
    scn[s].nbc = (int)e->bcl[s].bcs.size();
    scn[s].tile_base = tile_base;
    scn[s].staggered = fip_internal::staggered(e->sims[s]);
    tile_base += g.tiles_x * g.tiles_y;
    e->scn_tiles[s] = g.tiles_x * g.tiles_y;
    const ScnBCs& L = e->bcl[s];
    std::vector<int> stage, lim;
    fip_internal::bc_stages(L.bcs.data(), (int)L.bcs.size(), g.W, g.H, stage, lim);
    for (size_t i = 0; i < L.bcs.size(); i++) {
      const fip_bc& b = L.bcs[i];
      fip_ens_bc d{};
      d.type = b.type; d.side = b.side; d.from = b.from; d.to = b.to;
      d.z = b.z; d.q = b.q;
      if (!L.zt[i].empty()) {
        d.zoff = (int)series.size();
        d.zn = (int)L.zt[i].size();
        series.insert(series.end(), L.zt[i].begin(), L.zt[i].end());
        series.insert(series.end(), L.zv[i].begin(), L.zv[i].end());
      }
      if (!L.qt[i].empty()) {
        d.qoff = (int)series.size();
        d.qn = (int)L.qt[i].size();
        series.insert(series.end(), L.qt[i].begin(), L.qt[i].end());
        series.insert(series.end(), L.qv[i].begin(), L.qv[i].end());
      }
      const int gi = (int)bcs.size();
      bcs.push_back(d);
      if (stage[i] < 0) continue;
      if ((int)chunks.size() <= stage[i]) chunks.resize(stage[i] + 1);
      for (int p = b.from; p <= lim[i]; p += 256)
        chunks[stage[i]].push_back(fip_ens_chunk{s, gi, p, std::min(256, lim[i] - p + 1)});
    }
  }
  e->total_tiles = tile_base;
  if (!bcs.empty()) {
    EHIP(hipMalloc(&e->d_bcs, bcs.size() * sizeof(fip_ens_bc)));
    EHIP(hipMemcpy(e->d_bcs, bcs.data(), bcs.size() * sizeof(fip_ens_bc), hipMemcpyHostToDevice));
  }
  if (!series.empty()) {
    EHIP(hipMalloc(&e->d_series, series.size() * sizeof(float)));
    EHIP(hipMemcpy(e->d_series, series.data(), series.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  for (auto& c : chunks) {
    fip_ens_chunk* d = nullptr;
    if (!c.empty()) {
      EHIP(hipMalloc(&d, c.size() * sizeof(fip_ens_chunk)));
      EHIP(hipMemcpy(d, c.data(), c.size() * sizeof(fip_ens_chunk), hipMemcpyHostToDevice));
    }
    e->d_chunks.push_back(d);
    e->nchunks.push_back((int)c.size());
  }
  EHIP(hipMemcpy(e->d_scn, scn.data(), scn.size() * sizeof(fip_ens_scn), hipMemcpyHostToDevice));
  if (e->d_list) hipFree(e->d_list);
  e->d_list = nullptr;
  EHIP(hipMalloc(&e->d_list, 2 * (size_t)std::max(1, e->total_tiles) * sizeof(unsigned long long)));
  e->stats.tiles_total = e->total_tiles;
  e->built = true;
  return FIP_OK;
}

// Dense mode: the work items of the dense march, one per wavefront. Every scenario is cut
// into spans of one length; the tile columns of all span rows with the same first and
// last row (the first span row of every scenario, the interior ones, the last ones of
// scenarios of one height) are listed together, scenario by scenario and column by
// column, and cut into work-groups of wpb wavefronts, the last one padded with
// wavefronts past the grid. All wavefronts of a work-group then march the same rows, so
// the barriers of the march stay matched; neighbouring work-groups (which share halo
// columns) land on one XCD. The span is the shortest (at least 16 rows, span + 2 a
// multiple of the 6-row register ring) for which every work-group is resident at once in
// the fast variant, and 16 rows in the exact one, whose cost per row depends on the
// water: there the dispatcher balances short spans over many rounds (as in the single
// domain's dense mode).
int plan_dense(fip_ens* e) {
  int wpc = 0;
  hipError_t r = e->fast ? fip_ens_dense_waves_per_cu_fast(e->kvar ? 1 : 0, e->wpb, &wpc)
                         : fip_ens_dense_waves_per_cu_exact(e->kvar ? 1 : 0, e->wpb, &wpc);
  if (r != hipSuccess) return fail(FIP_E_HIP, std::string("occupancy query: ") + hipGetErrorString(r));
  if (wpc < 1) return fail(FIP_E_HIP, "the dense ensemble kernel cannot be resident");
  const long slots = (long)e->cus * (wpc / e->wpb * e->wpb);
  std::vector<int> tiles(e->n), heights(e->n);
  for (int s = 0; s < e->n; s++) {
    const fip_grid& g = fip_internal::grid(e->sims[s]);
    tiles[s] = g.tiles_x;
    heights[s] = g.H;
  }
  // wave-columns per row range (first, last), in order of first appearance
  auto ranges = [&](int span, std::vector<std::pair<std::pair<int, int>, std::vector<int2>>>& out) {
    out.clear();
    for (int s = 0; s < e->n; s++)
      for (int r0 = 0; r0 < heights[s]; r0 += span) {
        const std::pair<int, int> key(r0, std::min(r0 + span, heights[s]) - 1);
        size_t k = 0;
        while (k < out.size() && out[k].first != key) k++;
        if (k == out.size()) out.push_back({key, {}});
        for (int tx = 0; tx < tiles[s]; tx++) out[k].second.push_back(make_int2(s, tx));
      }
  };
  std::vector<std::pair<std::pair<int, int>, std::vector<int2>>> rg;
  auto waves = [&](int span) {  // the wavefronts of ranges(span), without listing them
    std::vector<std::pair<std::pair<int, int>, long>> cnt;
    for (int s = 0; s < e->n; s++)
      for (int r0 = 0; r0 < heights[s]; r0 += span) {
        const std::pair<int, int> key(r0, std::min(r0 + span, heights[s]) - 1);
        size_t k = 0;
        while (k < cnt.size() && cnt[k].first != key) k++;
        if (k == cnt.size()) cnt.push_back({key, 0});
        cnt[k].second += tiles[s];
      }
    long w = 0;
    for (auto& k : cnt) w += (k.second + e->wpb - 1) / e->wpb * e->wpb;
    return w;
  };
  const int hmax = *std::max_element(heights.begin(), heights.end());
  int span = 16;
  while (e->fast && waves(span) > slots && span < hmax) span += 6;
  ranges(span, rg);
  std::vector<int4> items;
  for (auto& k : rg) {
    const int first = k.first.first, last = k.first.second;
    for (const int2& c : k.second) items.push_back(make_int4(c.x, first, last, c.y));
    const int s0 = k.second.front().x;
    while (items.size() % e->wpb) items.push_back(make_int4(s0, first, last, tiles[s0]));
  }
  if (e->d_items) hipFree(e->d_items);
  e->d_items = nullptr;
  EHIP(hipMalloc(&e->d_items, items.size() * sizeof(int4)));
  EHIP(hipMemcpy(e->d_items, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice));
  e->nitems = (int)items.size();
  e->items_wpb = e->wpb;
  e->dense_span = span;
  e->stats.dense_span = span;
  return FIP_OK;
}

int launch_stencil(fip_ens* e, int par) {
  if (e->dense) {
    if (e->items_wpb != e->wpb) return fail(FIP_E_STATE, "dense work items not planned");
    if (e->timing) EHIP(hipEventRecord(e->ev0, e->stream));
    hipError_t r = e->fast ? fip_launch_ens_dense_fast(e->d_scn, e->d_items, e->nitems, e->kvar, e->wpb, par, e->stream)
                           : fip_launch_ens_dense_exact(e->d_scn, e->d_items, e->nitems, e->kvar, e->wpb, par, e->stream);
    if (r != hipSuccess) return fail(FIP_E_HIP, std::string("ensemble dense stencil launch: ") + hipGetErrorString(r));
    if (e->timing) {
      EHIP(hipEventRecord(e->ev1, e->stream));
      EHIP(hipEventSynchronize(e->ev1));
      float ms = 0.f;
      EHIP(hipEventElapsedTime(&ms, e->ev0, e->ev1));
      e->stats.stencil_ms += ms;
      e->stats.timed_launches++;
    }
    e->stats.stencil_launches++;
    return FIP_OK;
  }
  unsigned long long* list = e->d_list + (size_t)par * std::max(1, e->total_tiles);
  if (e->timing) EHIP(hipEventRecord(e->ev0, e->stream));
  const unsigned cap = (unsigned)e->total_tiles;
  const int persistent = e->bpc > 0 ? 1 : 0;
  const int blocks = persistent ? e->cus * e->bpc : std::max(1, (e->total_tiles + e->wpb - 1) / e->wpb);
  hipError_t r = e->fast ? fip_launch_ens_stencil_fast(e->d_scn, list, e->d_ctl, par, cap, e->R, e->kvar,
                                                        blocks, e->wpb, persistent, e->stream)
                         : fip_launch_ens_stencil_exact(e->d_scn, list, e->d_ctl, par, cap, e->R, e->kvar,
                                                         blocks, e->wpb, persistent, e->stream);
  if (r != hipSuccess) return fail(FIP_E_HIP, std::string("ensemble stencil launch: ") + hipGetErrorString(r));
  if (e->timing) {
    EHIP(hipEventRecord(e->ev1, e->stream));
    EHIP(hipEventSynchronize(e->ev1));
    float ms = 0.f;
    EHIP(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    e->stats.stencil_ms += ms;
    e->stats.timed_launches++;
  }
  e->stats.stencil_launches++;
  return FIP_OK;
}

// post-step control, then BC stages; the first stage's launch also schedules list q
int launch_post_apply(fip_ens* e, int mode, int par, int q) {
  hipError_t r = fip_launch_ens(0, e->d_scn, e->n, e->total_tiles, e->d_bcs, e->d_series, nullptr, 0,
                                e->d_ctl, nullptr, mode, par, -1, 0, 0, e->stream);
  if (r != hipSuccess) return fail(FIP_E_HIP, std::string("ensemble post launch: ") + hipGetErrorString(r));
  const int nst = std::max<int>(1, (int)e->nchunks.size());
  for (int st = 0; st < nst; st++) {
    const int nc = st < (int)e->nchunks.size() ? e->nchunks[st] : 0;
    const fip_ens_chunk* ch = st < (int)e->d_chunks.size() ? e->d_chunks[st] : nullptr;
    const int sched = st == 0 && !e->dense ? e->sched_blocks : 0;  // dense mode keeps no list
    if (nc + sched == 0) continue;
    unsigned long long* list = e->d_list + (size_t)q * std::max(1, e->total_tiles);
    r = fip_launch_ens(1, e->d_scn, e->n, e->total_tiles, e->d_bcs, e->d_series, ch, nc, e->d_ctl,
                       list, mode, par, st == 0 ? q : -1, e->dense ? 1 : 0, nc + sched, e->stream);
    if (r != hipSuccess) return fail(FIP_E_HIP, std::string("ensemble apply launch: ") + hipGetErrorString(r));
  }
  return FIP_OK;
}

int launch_flood(fip_ens* e) {
  hipError_t r = fip_launch_ens(2, e->d_scn, e->n, 0, nullptr, nullptr, nullptr, 0, nullptr, nullptr,
                                0, 0, -1, 0, 64, e->stream);
  if (r != hipSuccess) return fail(FIP_E_HIP, std::string("ensemble flood-plain launch: ") + hipGetErrorString(r));
  return FIP_OK;
}

int one_step(fip_ens* e, int par) {
  int rc = launch_stencil(e, par);
  if (!rc) rc = launch_post_apply(e, 1, par, par ^ 1);
  if (!rc && e->post_hook) rc = e->post_hook(e->stream, par);
  return rc;
}

// A run uses at most three chunk lengths (up to the next multiple of kStepsPerGraph, whole
// chunks, the rest), each at one or two parities.
constexpr size_t kMaxGraphs = 6;

// The graph of `steps` steps (1..kStepsPerGraph) from parity par, captured, instantiated
// and uploaded on first use; nullptr when that failed.
hipGraphExec_t chunk_graph(fip_ens* e, int par, int steps) {
  for (auto& c : e->graphs)
    if (c.par == par && c.steps == steps) return c.exec;
  hipGraph_t graph;
  if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
  int rc = FIP_OK;
  for (int k = 0; k < steps && rc == FIP_OK; k++) rc = one_step(e, par ^ (k & 1));
  if (rc == FIP_OK) rc = launch_flood(e);
  const hipError_t er = hipStreamEndCapture(e->stream, &graph);
  if (rc != FIP_OK || er != hipSuccess) return nullptr;
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ei != hipSuccess) return nullptr;
  if (hipGraphUpload(exec, e->stream) != hipSuccess || hipStreamSynchronize(e->stream) != hipSuccess) {
    hipGraphExecDestroy(exec);
    return nullptr;
  }
  if (e->graphs.size() == kMaxGraphs) {
    hipGraphExecDestroy(e->graphs.front().exec);
    e->graphs.erase(e->graphs.begin());
  }
  e->graphs.push_back({par, steps, exec});
  return exec;
}

int check_scn(fip_ens* e, int s) {
  if (!e) return fail(FIP_E_INVALID, "null argument");
  if (s < 0 || s >= e->n) return fail(FIP_E_INVALID, "scenario index out of range");
  return FIP_OK;
}

}  // namespace

namespace fip_internal {
void ens_set_post_hook(fip_ens* e, StepHook hook) {
  e->post_hook = std::move(hook);
  destroy_graphs(e);
}
const fip_ens_scn* ens_scn(const fip_ens* e) { return e->d_scn; }
int ens_count(const fip_ens* e) { return e->n; }
fip_sim* ens_member(const fip_ens* e, int s) { return e->sims[s]; }
hipStream_t ens_stream(const fip_ens* e) { return e->stream; }
}  // namespace fip_internal

extern "C" {

int fip_ens_create(const fip_desc* descs, int n, fip_ens** out) {
  if (!descs || !out) return fail(FIP_E_INVALID, "null argument");
  *out = nullptr;
  if (n < 1 || n > FIP_ENS_MAX)
    return fail(FIP_E_INVALID, "an ensemble holds 1 to " + std::to_string(FIP_ENS_MAX) + " scenarios");
  const bool kvar = descs[0].kst_var != nullptr;
  for (int s = 1; s < n; s++)
    if ((descs[s].kst_var != nullptr) != kvar)
      return fail(FIP_E_INVALID, "every scenario must have kst_var, or none");
  fip_ens* e = new fip_ens();
  e->n = n;
  e->dev = descs[0].device;
  e->R = descs[0].tile_rows == 0 ? 64 : descs[0].tile_rows;
  e->fast = descs[0].variant == FIP_FAST;
  e->dense = descs[0].mode == FIP_MODE_DENSE;
  e->kvar = kvar;
  e->wpb = e->dense && e->fast ? 16 : 4;
  e->bcl.resize(n);
  auto cleanup = [&](int rc) {
    fip_ens_destroy(e);
    return rc;
  };
  for (int s = 0; s < n; s++) {
    fip_desc d = descs[s];
    d.mode = descs[0].mode;
    d.variant = descs[0].variant;
    d.tile_rows = e->R;
    d.device = e->dev;
    fip_sim* sim = nullptr;
    int rc = fip_create(&d, &sim);
    if (rc) return cleanup(rc);
    e->sims.push_back(sim);
  }
  if (hipSetDevice(e->dev) != hipSuccess) return cleanup(fail(FIP_E_HIP, "hipSetDevice failed"));
  e->graph_failed = !fip_internal::graphs_enabled();
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(FIP_E_HIP, "hipStreamCreate failed"));
  if (hipMalloc(&e->d_scn, n * sizeof(fip_ens_scn)) != hipSuccess ||
      hipMalloc(&e->d_ctl, sizeof(fip_ens_ctl)) != hipSuccess ||
      hipMemset(e->d_ctl, 0, sizeof(fip_ens_ctl)) != hipSuccess)
    return cleanup(fail(FIP_E_HIP, "ensemble allocation failed"));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, e->dev) != hipSuccess)
    return cleanup(fail(FIP_E_HIP, "hipGetDeviceProperties failed"));
  const int cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  e->cus = cus;
  e->stencil_blocks = cus * 8;
  e->sched_blocks = cus;
  if (hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess)
    return cleanup(fail(FIP_E_HIP, "hipEventCreate failed"));
  *out = e;
  return FIP_OK;
}

int fip_ens_destroy(fip_ens* e) {
  if (!e) return FIP_OK;
  hipSetDevice(e->dev);
  if (e->stream) hipStreamSynchronize(e->stream);
  fip_internal::closure_forget(nullptr, e);
  e->post_hook = nullptr;
  destroy_graphs(e);
  free_tables(e);
  if (e->d_list) hipFree(e->d_list);
  if (e->d_items) hipFree(e->d_items);
  if (e->d_scn) hipFree(e->d_scn);
  if (e->d_ctl) hipFree(e->d_ctl);
  if (e->ev0) hipEventDestroy(e->ev0);
  if (e->ev1) hipEventDestroy(e->ev1);
  if (e->stream) hipStreamDestroy(e->stream);
  for (fip_sim* s : e->sims) fip_destroy(s);
  delete e;
  return FIP_OK;
}

int fip_ens_set_bcs(fip_ens* e, int s, const fip_bc* bcs, int n) {
  int rc = check_scn(e, s);
  if (rc) return rc;
  if (e->started) return fail(FIP_E_STATE, "BCs are fixed once the ensemble has started");
  if (n < 0 || n > FIP_MAX_BC) return fail(FIP_E_INVALID, "too many boundary conditions");
  if (n > 0 && !bcs) return fail(FIP_E_INVALID, "null argument");
  const fip_grid& g = fip_internal::grid(e->sims[s]);
  ScnBCs L;
  for (int i = 0; i < n; i++) {
    const fip_bc& b = bcs[i];
    if (b.type < 0 || b.type > 3) return fail(FIP_E_INVALID, "invalid boundary condition type");
    if (b.side < 0 || b.side > 3) return fail(FIP_E_INVALID, "invalid boundary condition side");
    const int ext = (b.side == FIP_SIDE_LEFT || b.side == FIP_SIDE_RIGHT) ? g.H : g.W;
    if (b.from < 0 || b.to >= ext) return fail(FIP_E_INVALID, "boundary condition range outside the side");
    for (int k = 1; k < b.zn; k++)
      if (!(b.zt[k] > b.zt[k - 1])) return fail(FIP_E_INVALID, "z series times must increase");
    for (int k = 1; k < b.qn; k++)
      if (!(b.qt[k] > b.qt[k - 1])) return fail(FIP_E_INVALID, "q series times must increase");
    fip_bc c = b;
    c.zt = c.zv = c.qt = c.qv = nullptr;
    L.bcs.push_back(c);
    L.zt.emplace_back(b.zn > 0 ? std::vector<float>(b.zt, b.zt + b.zn) : std::vector<float>());
    L.zv.emplace_back(b.zn > 0 ? std::vector<float>(b.zv, b.zv + b.zn) : std::vector<float>());
    L.qt.emplace_back(b.qn > 0 ? std::vector<float>(b.qt, b.qt + b.qn) : std::vector<float>());
    L.qv.emplace_back(b.qn > 0 ? std::vector<float>(b.qv, b.qv + b.qn) : std::vector<float>());
  }
  e->bcl[s] = std::move(L);
  e->built = false;
  return FIP_OK;
}

int fip_ens_set_run(fip_ens* e, const double* targets) {
  if (!e || !targets) return fail(FIP_E_INVALID, "null argument");
  for (int s = 0; s < e->n; s++) {
    int rc = fip_set_run(e->sims[s], targets[s], 0, nullptr, nullptr, nullptr, 0, nullptr, 0);
    if (rc) return rc;
  }
  e->run_set = true;
  return FIP_OK;
}

int fip_ens_set_launch(fip_ens* e, int waves_per_block, int blocks_per_cu) {
  if (!e) return fail(FIP_E_INVALID, "null argument");
  if (waves_per_block < 0 || waves_per_block > 16 || blocks_per_cu > 64)
    return fail(FIP_E_INVALID, "waves_per_block must be 0..16 and blocks_per_cu at most 64");
  e->wpb = waves_per_block != 0 ? waves_per_block : e->dense && e->fast ? 16 : 4;
  e->bpc = blocks_per_cu < 0 ? 0 : blocks_per_cu;
  destroy_graphs(e);
  return FIP_OK;
}

int fip_ens_tune(const fip_desc* descs, int n, const fip_bc* const* bcs, const int* nbcs,
                 int steps, fip_tune_result* out) {
  if (!descs || !out || !nbcs || (n > 0 && !bcs)) return fail(FIP_E_INVALID, "null argument");
  if (steps < 1) return fail(FIP_E_INVALID, "steps must be at least 1");
  static const int kRows[] = {16, 32, 64};
  static const int kWaves[] = {1, 2, 4, 8};
  static const int kBlocksPerCu[] = {0, 4, 8, 16};
  static const int kDenseWaves[] = {2, 4, 8, 16};
  const bool dense = n > 0 && descs[0].mode == FIP_MODE_DENSE;
  std::vector<int> rows(kRows, kRows + 3), waves(kWaves, kWaves + 4), bpcs(kBlocksPerCu, kBlocksPerCu + 4);
  if (dense) {  // the dense march uses neither the tile rows nor persistent work-groups
    rows.assign(1, descs[0].tile_rows == 0 ? 64 : descs[0].tile_rows);
    waves.assign(kDenseWaves, kDenseWaves + 4);
    bpcs.assign(1, 0);
  }
  std::vector<fip_desc> dd(descs, descs + n);
  fip_tune_result best{};
  best.stencil_ms = 1e30;
  std::vector<double> far(n, 1e300);
  for (int R : rows) {
    for (auto& d : dd) d.tile_rows = R;
    fip_ens* e = nullptr;
    int rc = fip_ens_create(dd.data(), n, &e);
    for (int s = 0; s < n && !rc; s++) rc = fip_ens_set_bcs(e, s, bcs[s], nbcs[s]);
    if (!rc) rc = fip_ens_set_run(e, far.data());
    for (int w : waves)
      for (int bpc : bpcs) {
        if (rc) break;
        int act = 0;
        long long done = 0;
        rc = fip_ens_set_launch(e, w, bpc);
        if (!rc) rc = fip_ens_set_timing(e, 0);
        if (!rc) rc = fip_ens_simulate(e, steps, &act, &done);
        const fip_stats before = e ? e->stats : fip_stats{};
        if (!rc) rc = fip_ens_set_timing(e, 1);
        if (!rc) rc = fip_ens_simulate(e, steps, &act, &done);
        if (rc) break;
        const long long k = e->stats.timed_launches - before.timed_launches;
        const double ms = k > 0 ? (e->stats.stencil_ms - before.stencil_ms) / (double)k : 1e30;
        best.tried++;
        if (ms < best.stencil_ms) {
          best.stencil_ms = ms;
          best.tile_rows = R;
          best.waves_per_block = w;
          best.blocks_per_cu = bpc;
        }
      }
    if (e) fip_ens_destroy(e);
    if (rc) return rc;
  }
  *out = best;
  return FIP_OK;
}

int fip_ens_prepare(fip_ens* e, long long nsteps) {
  if (!e) return fail(FIP_E_INVALID, "null argument");
  if (!e->run_set) return fail(FIP_E_STATE, "call fip_ens_set_run first");
  if (nsteps < 0) return fail(FIP_E_INVALID, "nsteps must not be negative");
  EHIP(hipSetDevice(e->dev));
  int rc;
  if (!e->built && (rc = build(e))) return rc;
  if (e->dense && e->items_wpb != e->wpb && (rc = plan_dense(e))) return rc;
  if (e->timing || e->graph_failed) return FIP_OK;
  // the chunks fip_ens_simulate cuts the steps into (chunk_graph returns the kept ones)
  for (long long it = e->iter, left = nsteps; left > 0;) {
    const long long k = std::min<long long>(kStepsPerGraph - (it % kStepsPerGraph), left);
    if (!chunk_graph(e, (int)(it & 1), (int)k)) {
      e->graph_failed = true;
      hipGetLastError();
      break;
    }
    it += k;
    left -= k;
  }
  return FIP_OK;
}

int fip_ens_simulate(fip_ens* e, long long nsteps, int* active, long long* steps_done) {
  if (!e) return fail(FIP_E_INVALID, "null argument");
  if (steps_done) *steps_done = 0;
  if (!e->run_set) return fail(FIP_E_STATE, "call fip_ens_set_run first");
  EHIP(hipSetDevice(e->dev));
  int rc;
  if (!e->built && (rc = build(e))) return rc;
  if (e->dense && e->items_wpb != e->wpb && (rc = plan_dense(e))) return rc;
  if (!e->started) {
    if ((rc = launch_post_apply(e, 0, (int)(e->iter & 1), (int)(e->iter & 1)))) return rc;
    e->started = true;
  }
  std::vector<int> fin(e->n, 0);
  auto count_active = [&](int* out) -> int {
    EHIP(hipMemcpyAsync(fin.data(), (char*)e->d_ctl + offsetof(fip_ens_ctl, finished), sizeof(int) * e->n,
                        hipMemcpyDeviceToHost, e->stream));
    EHIP(hipStreamSynchronize(e->stream));
    int a = 0;
    for (int v : fin) a += v ? 0 : 1;
    *out = a;
    return FIP_OK;
  };
  int act = 0;
  if ((rc = count_active(&act))) return rc;
  long long done = 0;
  while (done < nsteps && act > 0) {
    const int par = (int)(e->iter & 1);
    const long long to_flood = kStepsPerGraph - (e->iter % kStepsPerGraph);
    const long long k = std::min<long long>(to_flood, nsteps - done);
    hipGraphExec_t gx = nullptr;
    if (!e->timing && !e->graph_failed) {
      gx = chunk_graph(e, par, (int)k);
      if (!gx) {
        e->graph_failed = true;
        hipGetLastError();
      }
    }
    if (gx) {
      EHIP(hipGraphLaunch(gx, e->stream));
      e->stats.stencil_launches += k;
    } else {
      for (long long i = 0; i < k; i++)
        if ((rc = one_step(e, (int)((e->iter + i) & 1)))) return rc;
      if ((rc = launch_flood(e))) return rc;
    }
    e->iter += k;
    done += k;
    e->stats.steps += k;
    if ((rc = count_active(&act))) return rc;
  }
  if (e->dense) {  // every tile of every running scenario
    long long t = 0;
    for (int s = 0; s < e->n; s++) t += fin[s] ? 0 : e->scn_tiles[s];
    e->stats.tiles_computed = t;
  } else {
    unsigned cnt[2];
    EHIP(hipMemcpy(cnt, (char*)e->d_ctl + offsetof(fip_ens_ctl, list_count), sizeof cnt, hipMemcpyDeviceToHost));
    e->stats.tiles_computed = cnt[e->iter & 1];  // the list of the next step
  }
  for (fip_sim* s : e->sims)
    if ((rc = fip_internal::refresh(s))) return rc;
  if (active) *active = act;
  if (steps_done) *steps_done = done;
  return FIP_OK;
}

int fip_ens_get_time(fip_ens* e, int s, double* duration, float* dt, int* iteration) {
  int rc = check_scn(e, s);
  if (rc) return rc;
  EHIP(hipSetDevice(e->dev));
  EHIP(hipStreamSynchronize(e->stream));
  return fip_get_time(e->sims[s], duration, dt, iteration);
}

int fip_ens_download(fip_ens* e, int s, int field, int side, float* dst) {
  int rc = check_scn(e, s);
  if (rc) return rc;
  EHIP(hipSetDevice(e->dev));
  EHIP(hipStreamSynchronize(e->stream));
  if ((rc = fip_internal::refresh(e->sims[s]))) return rc;
  return fip_download(e->sims[s], field, side, dst);
}

int fip_ens_set_timing(fip_ens* e, int enable) {
  if (!e) return fail(FIP_E_INVALID, "null argument");
  e->timing = enable != 0;
  return FIP_OK;
}

int fip_ens_get_stats(fip_ens* e, fip_stats* out) {
  if (!e || !out) return fail(FIP_E_INVALID, "null argument");
  *out = e->stats;
  return FIP_OK;
}

}  // extern "C"
