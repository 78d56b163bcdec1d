// libfiphip: host side of the learned-closure C ABI (include/fip_closure.h): parameters
// and Adam state, the training-sample store, the training and evaluation loops, and the
// inference hook that corrects a coarse simulation (or ensemble) after every step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fip_closure.h"
#include "closure.h"
#include "fip_internal.h"

using fip_internal::fail;

static_assert(CL_ACT_GELU == FIP_CLOSURE_ACT_GELU && CL_ACT_HSILU == FIP_CLOSURE_ACT_HSILU,
              "activation codes of the C ABI and the kernels");

#define CHIP(expr)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(FIP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
  } while (0)

namespace {

constexpr float kBound = 0.1f;   // |c| <= 0.1 |q|
constexpr int kWgradBlocks = 768;  // 3 work-groups of 42.5 KB LDS per CU

// splitmix64 stream with uniform and normal draws (Box-Muller)
struct Rng {
  uint64_t s;
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

template <class T>
int dalloc(T** p, size_t n) {
  if (*p) hipFree(*p);
  *p = nullptr;
  if (n == 0) return FIP_OK;
  CHIP(hipMalloc(p, n * sizeof(T)));
  return FIP_OK;
}

// Activations of one batch shape (N images of H x W), NHWC bf16 with 32 channels.
struct Work {
  int N = 0, H = 0, W = 0;
  u16* x0 = nullptr;
  u16* z[3] = {};
  u16* a[3] = {};
  u16* dz[4] = {};
  float* o = nullptr;
  float* qp = nullptr;
  float* qt = nullptr;
  uint8_t* mask = nullptr;
  int* origin = nullptr;
  double* lpart = nullptr;
  ~Work() { release(); }
  void release() {
    auto f = [](void* p) {
      if (p) hipFree(p);
    };
    f(x0);
    for (auto* p : z) f(p);
    for (auto* p : a) f(p);
    for (auto* p : dz) f(p);
    f(o); f(qp); f(qt); f(mask); f(origin); f(lpart);
    x0 = nullptr; o = qp = qt = nullptr; mask = nullptr; origin = nullptr; lpart = nullptr;
    for (auto& p : z) p = nullptr;
    for (auto& p : a) p = nullptr;
    for (auto& p : dz) p = nullptr;
    N = H = W = 0;
  }
  long pixels() const { return (long)N * H * W; }
  // forward tensors always; backward and target tensors when `train`
  int ensure(int n, int h, int w, bool train) {
    if (n == N && h == H && w == W && (!train || dz[0])) return FIP_OK;
    release();
    N = n; H = h; W = w;
    const size_t act = (size_t)n * h * w * CL_C;
    int rc = dalloc(&x0, act);
    for (int k = 0; k < 3 && !rc; k++) rc = dalloc(&z[k], act) | dalloc(&a[k], act);
    if (!rc) rc = dalloc(&o, (size_t)n * h * w * 2);
    if (!rc) rc = dalloc(&lpart, (size_t)((pixels() + 255) / 256) * 3);
    if (train) {
      for (int k = 0; k < 4 && !rc; k++) rc = dalloc(&dz[k], act);
      if (!rc) rc = dalloc(&qp, (size_t)n * h * w * 2);
      if (!rc) rc = dalloc(&qt, (size_t)n * h * w * 2);
      if (!rc) rc = dalloc(&mask, (size_t)n * h * w * 2);
      if (!rc) rc = dalloc(&origin, (size_t)n * 3);
    }
    return rc;
  }
};

struct Attachment {
  fip_sim* sim = nullptr;   // single domain, or
  fip_ens* ens = nullptr;   // an ensemble
  float* bstd = nullptr;
  float* zref = nullptr;    // device, one per scenario (ensemble)
  float zref_h = 0.f;
  ClMass* ms = nullptr;     // one per scenario
  double* mpart = nullptr;  // mass-balance terms per correction work-group, [scenario][nparts][2]
  ClDom* dom = nullptr;     // one per scenario
  int4* items = nullptr;    // work items of the row-march network kernel (cl_fused_plan)
  int nitems = 0;
  int wmax = 0, hmax = 0;   // the largest scenario size
  float* o = nullptr;       // network outputs, each scenario's [H][W][2] image in turn
  unsigned* part = nullptr; // saturated faces per correction work-group, [scenario][nparts]
  int nparts = 0;
  float sat = 2.65f;
  double limit = 1e-4;
  ~Attachment() {
    if (o) hipFree(o);
    if (part) hipFree(part);
    if (bstd) hipFree(bstd);
    if (zref) hipFree(zref);
    if (ms) hipFree(ms);
    if (mpart) hipFree(mpart);
    if (dom) hipFree(dom);
    if (items) hipFree(items);
  }
};

int device_cus(int dev) {
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) n = 256;
  return n;
}

// the row march's work items for domains of the given sizes, uploaded to *items; `extra`
// compute units are left to the kernel's mass-check work-groups
int upload_plan(const std::vector<int>& W, const std::vector<int>& H, int dev, int4** items, int* nitems,
                int extra = 0) {
  const std::vector<int4> it =
      cl_fused_plan(W.data(), H.data(), (int)W.size(), std::max(1, device_cus(dev) - extra));
  int rc = dalloc(items, it.size());
  if (rc) return rc;
  CHIP(hipMemcpy(*items, it.data(), it.size() * sizeof(int4), hipMemcpyHostToDevice));
  *nitems = (int)it.size();
  return FIP_OK;
}

}  // namespace

struct fip_closure {
  int dev = 0;
  hipStream_t stream = nullptr;
  float* w = nullptr;   // fp32 master parameters [layer][CL_LAYER_PARAMS]
  float* m = nullptr;
  float* v = nullptr;
  float* g = nullptr;
  u16* wf = nullptr;    // packed forward weights [layer][CL_PACKED]
  u16* wb = nullptr;    // packed backward-data weights
  u16* wl1 = nullptr;   // layer 0 in the 8-channel packing of the fused inference kernel
  float* part = nullptr;
  int t = 0;            // Adam steps taken
  int fused_form = 2;   // launch form of the fused inference kernel (cl_fused_launch)
  int act = CL_ACT_GELU;  // the network's activation
  // training samples
  int Wc = 0, Hc = 0, cap = 0, ns = 0;
  u16* s_feat = nullptr;     // [cap][Hc][Wc][8]
  float* s_qp = nullptr;     // [cap][Hc][Wc][2]
  float* s_qt = nullptr;
  uint8_t* s_mask = nullptr;
  Work train;
  Work eval;
  Work host;
  std::vector<std::unique_ptr<Attachment>> att;
};

namespace {

const float* bias_of(const fip_closure* cl, int l) { return cl->w + (size_t)l * CL_LAYER_PARAMS + CL_TAPS * CL_C * CL_C; }

int pack(fip_closure* cl) {
  CHIP(cl_pack_launch(cl->w, CL_LAYERS, cl->wf, cl->wb, cl->stream));
  CHIP(cl_pack_l1_launch(cl->w, cl->wl1, cl->stream));
  return FIP_OK;
}

// forward pass of ws (x0 already filled); ctl/par: skip when that domain's step was skipped
int forward(fip_closure* cl, Work& ws, const fip_ctl* ctl, int par, hipStream_t st) {
  const u16* in = ws.x0;
  for (int l = 0; l < 3; l++) {
    CHIP(cl_launch_conv(EPI_ACT, cl->act, in, cl->wf + (size_t)l * CL_PACKED, bias_of(cl, l), nullptr, ws.z[l],
                        ws.a[l], nullptr, ws.N, ws.H, ws.W, ctl, par, st));
    in = ws.a[l];
  }
  CHIP(cl_launch_conv(EPI_OUT, cl->act, in, cl->wf + 3 * (size_t)CL_PACKED, bias_of(cl, 3), nullptr, nullptr,
                      nullptr, ws.o, ws.N, ws.H, ws.W, ctl, par, st));
  return FIP_OK;
}

// loss + backward into cl->g; returns the partial sums in ws.lpart
int backward(fip_closure* cl, Work& ws, float bound, hipStream_t st) {
  const long npix = ws.pixels();
  CHIP(cl_loss_launch(ws.o, ws.qp, ws.qt, ws.mask, npix, bound, 1.0f / (2.0f * (float)npix), ws.dz[3],
                      ws.lpart, st));
  const int nb = (int)std::min<long>(kWgradBlocks, std::max<long>(1, (long)ws.N * ((ws.H + 3) / 4) * ((ws.W + 63) / 64)));
  const u16* xin[4] = {ws.x0, ws.a[0], ws.a[1], ws.a[2]};
  for (int l = 3; l >= 0; l--) {
    CHIP(cl_wgrad_launch(xin[l], ws.dz[l], cl->part, ws.N, ws.H, ws.W, nb, st));
    CHIP(cl_reduce_launch(cl->part, nb, CL_LAYER_PARAMS, cl->g + (size_t)l * CL_LAYER_PARAMS, st));
    if (l > 0)
      CHIP(cl_launch_conv(EPI_DACT, cl->act, ws.dz[l], cl->wb + (size_t)l * CL_PACKED, nullptr, ws.z[l - 1],
                          ws.dz[l - 1], nullptr, nullptr, ws.N, ws.H, ws.W, nullptr, 0, st));
  }
  return FIP_OK;
}

int read_loss(Work& ws, hipStream_t st, double* sum_r, double* sum_m, double* sum_p) {
  const int nb = (int)((ws.pixels() + 255) / 256);
  std::vector<double> h((size_t)nb * 3);
  CHIP(hipMemcpyAsync(h.data(), ws.lpart, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  CHIP(hipStreamSynchronize(st));
  double r = 0, m = 0, p = 0;
  for (int b = 0; b < nb; b++) {
    r += h[(size_t)b * 3];
    m += h[(size_t)b * 3 + 1];
    p += h[(size_t)b * 3 + 2];
  }
  *sum_r = r;
  *sum_m = m;
  *sum_p = p;
  return FIP_OK;
}

int init_params(fip_closure* cl, uint64_t seed) {
  std::vector<float> w((size_t)CL_LAYERS * CL_LAYER_PARAMS, 0.f);
  Rng rng{seed};
  for (int l = 0; l < CL_LAYERS; l++) {
    const int cin = l == 0 ? CL_FEAT : CL_C, cout = l == CL_LAYERS - 1 ? CL_OUT : CL_C;
    const double sd = std::sqrt(2.0 / (9.0 * cin)) * (l == CL_LAYERS - 1 ? 0.1 : 1.0);
    for (int t = 0; t < CL_TAPS; t++)
      for (int ci = 0; ci < cin; ci++)
        for (int co = 0; co < cout; co++)
          w[(size_t)l * CL_LAYER_PARAMS + ((size_t)t * CL_C + ci) * CL_C + co] = (float)(sd * rng.normal());
  }
  CHIP(hipMemcpy(cl->w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
  return pack(cl);
}

Attachment* find_att(fip_closure* cl, fip_sim* sim, fip_ens* ens) {
  for (auto& a : cl->att)
    if ((sim && a->sim == sim) || (ens && a->ens == ens)) return a.get();
  return nullptr;
}

// all closures that hold an attachment of a simulation (to detach it)
std::vector<fip_closure*>& registry() {
  static std::vector<fip_closure*> r;
  return r;
}

}  // namespace

namespace fip_internal {
void closure_forget(const fip_sim* s, const fip_ens* e) {
  for (fip_closure* cl : registry())
    for (auto it = cl->att.begin(); it != cl->att.end();) {
      if ((s && (*it)->sim == s) || (e && (*it)->ens == e)) {
        hipSetDevice(cl->dev);
        hipDeviceSynchronize();
        it = cl->att.erase(it);
      } else {
        ++it;
      }
    }
}
}  // namespace fip_internal

extern "C" {

int fip_restrict(fip_sim* fine, fip_sim* coarse, int F) {
  if (!fine || !coarse) return fail(FIP_E_INVALID, "null argument");
  const fip_grid& f = fip_internal::grid(fine);
  const fip_grid& c = fip_internal::grid(coarse);
  if (F < 1 || c.W * F != f.W || c.H * F != f.H)
    return fail(FIP_E_INVALID, "coarse grid must be exactly W/factor x H/factor of the fine grid");
  CHIP(hipSetDevice(fip_internal::device(fine)));
  CHIP(hipDeviceSynchronize());
  CHIP(cl_restrict_launch(f, fip_internal::cur_side(fine), c, fip_internal::cur_side(coarse), F, nullptr));
  CHIP(hipDeviceSynchronize());
  return FIP_OK;
}

// Scratch of fip_closure_forward_noalloc: four bf16 images of 32 channels for up to
// FIP_CL_NOALLOC_MAX_PIXELS pixels (64 MiB): the network input, one pre-activation image
// that the three hidden layers overwrite in turn (the forward pass never reads it back),
// and two activation images used alternately. Every pixel and channel of an image is
// written before it is read, so nothing needs clearing between calls. Each device holds
// its own copy; its address is looked up once per device, at fip_closure_create.
namespace {
constexpr size_t kNoAllocImage = (size_t)FIP_CL_NOALLOC_MAX_PIXELS * CL_C;
constexpr int kNoAllocDevices = 64;
void* g_scratch_ptr[kNoAllocDevices] = {};
std::mutex g_scratch_mu;
}  // namespace
static __device__ __align__(256) u16 scratch[4 * kNoAllocImage];

namespace {
// With `device` current: the device address of its copy of `scratch`.
int noalloc_setup(int device) {
  if (device < 0 || device >= kNoAllocDevices) return fail(FIP_E_INVALID, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return FIP_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(FIP_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return FIP_OK;
}
}  // namespace

int fip_closure_create(int device, unsigned long long seed, int Wc, int Hc, int max_samples, fip_closure** out) {
  if (!out) return fail(FIP_E_INVALID, "null argument");
  *out = nullptr;
  if (Wc < 8 || Hc < 8 || max_samples < 0) return fail(FIP_E_INVALID, "coarse grid too small");
  fip_closure* cl = new fip_closure();
  auto cleanup = [&](int rc) {
    fip_closure_destroy(cl);
    return rc;
  };
  cl->dev = device;
  cl->Wc = Wc;
  cl->Hc = Hc;
  cl->cap = max_samples;
  if (hipSetDevice(device) != hipSuccess) return cleanup(fail(FIP_E_HIP, "hipSetDevice failed (no GPU?)"));
  if (hipStreamCreateWithFlags(&cl->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(FIP_E_HIP, "hipStreamCreate failed"));
  const size_t np = (size_t)CL_LAYERS * CL_LAYER_PARAMS;
  int rc = noalloc_setup(device);
  if (rc) return cleanup(rc);
  rc = dalloc(&cl->w, np);
  if (!rc) rc = dalloc(&cl->m, np);
  if (!rc) rc = dalloc(&cl->v, np);
  if (!rc) rc = dalloc(&cl->g, np);
  if (!rc) rc = dalloc(&cl->wf, (size_t)CL_LAYERS * CL_PACKED);
  if (!rc) rc = dalloc(&cl->wb, (size_t)CL_LAYERS * CL_PACKED);
  if (!rc) rc = dalloc(&cl->wl1, (size_t)CL_PACKED_L1);
  if (!rc) rc = dalloc(&cl->part, (size_t)kWgradBlocks * CL_LAYER_PARAMS);
  const size_t spx = (size_t)Wc * Hc;
  if (!rc) rc = dalloc(&cl->s_feat, (size_t)max_samples * spx * 8);
  if (!rc) rc = dalloc(&cl->s_qp, (size_t)max_samples * spx * 2);
  if (!rc) rc = dalloc(&cl->s_qt, (size_t)max_samples * spx * 2);
  if (!rc) rc = dalloc(&cl->s_mask, (size_t)max_samples * spx * 2);
  if (rc) return cleanup(rc);
  if (hipMemset(cl->m, 0, np * sizeof(float)) != hipSuccess || hipMemset(cl->v, 0, np * sizeof(float)) != hipSuccess ||
      hipMemset(cl->wf, 0, (size_t)CL_LAYERS * CL_PACKED * 2) != hipSuccess ||
      hipMemset(cl->wb, 0, (size_t)CL_LAYERS * CL_PACKED * 2) != hipSuccess ||
      hipMemset(cl->wl1, 0, (size_t)CL_PACKED_L1 * 2) != hipSuccess)
    return cleanup(fail(FIP_E_HIP, "hipMemset failed"));
  if ((rc = init_params(cl, seed))) return cleanup(rc);
  if (hipStreamSynchronize(cl->stream) != hipSuccess) return cleanup(fail(FIP_E_HIP, "sync failed"));
  registry().push_back(cl);
  *out = cl;
  return FIP_OK;
}

int fip_closure_destroy(fip_closure* cl) {
  if (!cl) return FIP_OK;
  hipSetDevice(cl->dev);
  for (auto& a : cl->att) {
    if (a->sim) fip_internal::set_post_hook(a->sim, nullptr);
    if (a->ens) fip_internal::ens_set_post_hook(a->ens, nullptr);
  }
  if (cl->stream) hipStreamSynchronize(cl->stream);
  cl->att.clear();
  for (void* p : {(void*)cl->w, (void*)cl->m, (void*)cl->v, (void*)cl->g, (void*)cl->wf, (void*)cl->wb, (void*)cl->wl1,
                  (void*)cl->part, (void*)cl->s_feat, (void*)cl->s_qp, (void*)cl->s_qt, (void*)cl->s_mask})
    if (p) hipFree(p);
  cl->train.release();
  cl->eval.release();
  cl->host.release();
  if (cl->stream) hipStreamDestroy(cl->stream);
  auto& r = registry();
  r.erase(std::remove(r.begin(), r.end(), cl), r.end());
  delete cl;
  return FIP_OK;
}

int fip_closure_get_params(fip_closure* cl, float* params) {
  if (!cl || !params) return fail(FIP_E_INVALID, "null argument");
  CHIP(hipSetDevice(cl->dev));
  CHIP(hipStreamSynchronize(cl->stream));
  CHIP(hipMemcpy(params, cl->w, sizeof(float) * CL_LAYERS * CL_LAYER_PARAMS, hipMemcpyDeviceToHost));
  return FIP_OK;
}

int fip_closure_set_params(fip_closure* cl, const float* params) {
  if (!cl || !params) return fail(FIP_E_INVALID, "null argument");
  CHIP(hipSetDevice(cl->dev));
  CHIP(hipStreamSynchronize(cl->stream));
  const size_t np = (size_t)CL_LAYERS * CL_LAYER_PARAMS;
  CHIP(hipMemcpy(cl->w, params, sizeof(float) * np, hipMemcpyHostToDevice));
  // new parameters start a new optimisation: Adam's moments and step count from zero
  CHIP(hipMemset(cl->m, 0, sizeof(float) * np));
  CHIP(hipMemset(cl->v, 0, sizeof(float) * np));
  cl->t = 0;
  int rc = pack(cl);
  if (rc) return rc;
  CHIP(hipStreamSynchronize(cl->stream));
  return FIP_OK;
}

int fip_closure_add_sample(fip_closure* cl, fip_sim* coarse, fip_sim* fine, int F, const float* bed_std,
                           float zref) {
  if (!cl || !coarse || !fine || !bed_std) return fail(FIP_E_INVALID, "null argument");
  if (cl->ns >= cl->cap) return fail(FIP_E_INVALID, "sample store is full");
  const fip_grid& c = fip_internal::grid(coarse);
  const fip_grid& f = fip_internal::grid(fine);
  if (c.W != cl->Wc || c.H != cl->Hc) return fail(FIP_E_INVALID, "coarse grid size differs from the closure's");
  if (c.W * F != f.W || c.H * F != f.H) return fail(FIP_E_INVALID, "fine grid must be factor x the coarse grid");
  CHIP(hipSetDevice(cl->dev));
  CHIP(hipDeviceSynchronize());
  const int cur = fip_internal::cur_side(coarse), oth = cur ^ 1;
  const size_t spx = (size_t)c.W * c.H;
  float* bstd = nullptr;
  CHIP(hipMalloc(&bstd, spx * sizeof(float)));
  CHIP(hipMemcpy(bstd, bed_std, spx * sizeof(float), hipMemcpyHostToDevice));
  ClFeat fe{c.h[oth], c.qx[oth], c.qy[oth], c.bed, c.kst, bstd, c.kst_uniform, zref, c.W, c.H, c.P};
  const size_t s = (size_t)cl->ns;
  hipError_t e = cl_features_launch(fe, cl->s_feat + s * spx * 8, 8, nullptr, 0, cl->stream);
  if (e == hipSuccess) e = cl_flux_launch(c, cur, cl->s_qp + s * spx * 2, cl->stream);
  if (e == hipSuccess)
    e = cl_restrict_flux_launch(f, fip_internal::cur_side(fine), F, c.W, c.H, cl->s_qt + s * spx * 2,
                                cl->s_mask + s * spx * 2, cl->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(cl->stream);
  hipFree(bstd);
  if (e != hipSuccess) return fail(FIP_E_HIP, std::string("add_sample: ") + hipGetErrorString(e));
  cl->ns++;
  return FIP_OK;
}

int fip_closure_num_samples(fip_closure* cl, int* n) {
  if (!cl || !n) return fail(FIP_E_INVALID, "null argument");
  *n = cl->ns;
  return FIP_OK;
}

namespace {
// One Adam step (beta 0.9 / 0.999, eps 1e-8) of the fp32 master weights from the gradient
// in cl->g, then the repack of the bf16 fragments.
int adam_update(fip_closure* cl, float lr) {
  cl->t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)cl->t), c2 = 1.f - std::pow(b2, (float)cl->t);
  CHIP(cl_adam_launch(cl->w, cl->m, cl->v, cl->g, CL_LAYERS, lr, b1, b2, 1e-8f, c1, c2, cl->stream));
  return pack(cl);
}
}  // namespace

int fip_closure_adam_step(fip_closure* cl, const float* grad, float lr) {
  if (!cl || !grad) return fail(FIP_E_INVALID, "null argument");
  CHIP(hipSetDevice(cl->dev));
  CHIP(hipMemcpyAsync(cl->g, grad, sizeof(float) * CL_LAYERS * CL_LAYER_PARAMS, hipMemcpyHostToDevice,
                      cl->stream));
  int rc = adam_update(cl, lr);
  if (rc) return rc;
  CHIP(hipStreamSynchronize(cl->stream));
  return FIP_OK;
}

int fip_closure_train(fip_closure* cl, int steps, int batch, int patch, float lr, int first, int n,
                      unsigned long long seed, float* losses) {
  if (!cl) return fail(FIP_E_INVALID, "null argument");
  if (n < 1 || first < 0 || first + n > cl->ns) return fail(FIP_E_INVALID, "sample range outside the store");
  if (patch < 8 || patch > std::min(cl->Wc, cl->Hc) || batch < 1)
    return fail(FIP_E_INVALID, "patch must fit the coarse grid");
  if (steps < 0) return fail(FIP_E_INVALID, "steps must not be negative");
  if (steps == 0) return FIP_OK;
  CHIP(hipSetDevice(cl->dev));
  int rc = cl->train.ensure(batch, patch, patch, true);
  if (rc) return rc;
  // every step's batch origins are drawn up front and uploaded once, and every step's loss
  // stays on the device until the end: the steps run back to back with no host round trip
  Rng rng{seed};
  std::vector<int> org((size_t)steps * batch * 3);
  for (int it = 0; it < steps; it++)
    for (int b = 0; b < batch; b++) {
      int* o = &org[((size_t)it * batch + b) * 3];
      o[0] = first + (int)(rng.uniform() * n) % n;
      o[1] = (int)(rng.uniform() * (cl->Wc - patch + 1)) % (cl->Wc - patch + 1);
      o[2] = (int)(rng.uniform() * (cl->Hc - patch + 1)) % (cl->Hc - patch + 1);
    }
  int* d_org = nullptr;
  float* d_loss = nullptr;
  if ((rc = dalloc(&d_org, org.size()))) return rc;
  if ((rc = dalloc(&d_loss, (size_t)steps))) {
    hipFree(d_org);
    return rc;
  }
  const size_t spx = (size_t)cl->Wc * cl->Hc;
  Work& ws = cl->train;
  const int nb = (int)((ws.pixels() + 255) / 256);
  hipError_t e = hipMemcpyAsync(d_org, org.data(), org.size() * sizeof(int), hipMemcpyHostToDevice, cl->stream);
  for (int it = 0; it < steps && e == hipSuccess && !rc; it++) {
    const int* o = d_org + (size_t)it * batch * 3;
    e = cl_gather_launch(cl->s_feat, (long)spx * 8, cl->Wc, o, patch, patch, ws.x0, batch, cl->stream);
    if (e == hipSuccess)
      e = cl_gather_targets_launch(cl->s_qp, cl->s_qt, cl->s_mask, (long)spx, cl->Wc, o, patch, patch, ws.qp,
                                   ws.qt, ws.mask, batch, cl->stream);
    if (e != hipSuccess) break;
    if ((rc = forward(cl, ws, nullptr, 0, cl->stream))) break;
    if ((rc = backward(cl, ws, kBound, cl->stream))) break;
    e = cl_loss_sum_launch(ws.lpart, nb, d_loss + it, cl->stream);
    if (e == hipSuccess) rc = adam_update(cl, lr);
  }
  if (e == hipSuccess && !rc && losses)
    e = hipMemcpyAsync(losses, d_loss, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost, cl->stream);
  const hipError_t es = hipStreamSynchronize(cl->stream);
  hipFree(d_org);
  hipFree(d_loss);
  if (rc) return rc;
  if (e != hipSuccess || es != hipSuccess)
    return fail(FIP_E_HIP, std::string("train: ") + hipGetErrorString(e != hipSuccess ? e : es));
  return FIP_OK;
}

int fip_closure_eval(fip_closure* cl, int first, int n, double* mse_plain, double* mse_learned) {
  if (!cl) return fail(FIP_E_INVALID, "null argument");
  if (n < 1 || first < 0 || first + n > cl->ns) return fail(FIP_E_INVALID, "sample range outside the store");
  CHIP(hipSetDevice(cl->dev));
  Work& ws = cl->eval;
  int rc = ws.ensure(1, cl->Hc, cl->Wc, true);
  if (rc) return rc;
  const size_t spx = (size_t)cl->Wc * cl->Hc;
  double R = 0, M = 0, Pp = 0;
  for (int s = first; s < first + n; s++) {
    const int org[3] = {s, 0, 0};
    CHIP(hipMemcpyAsync(ws.origin, org, sizeof org, hipMemcpyHostToDevice, cl->stream));
    CHIP(cl_gather_launch(cl->s_feat, (long)spx * 8, cl->Wc, ws.origin, cl->Wc, cl->Hc, ws.x0, 1, cl->stream));
    CHIP(cl_gather_targets_launch(cl->s_qp, cl->s_qt, cl->s_mask, (long)spx, cl->Wc, ws.origin, cl->Wc, cl->Hc,
                                  ws.qp, ws.qt, ws.mask, 1, cl->stream));
    if ((rc = forward(cl, ws, nullptr, 0, cl->stream))) return rc;
    CHIP(cl_loss_launch(ws.o, ws.qp, ws.qt, ws.mask, ws.pixels(), kBound, 1.f, ws.dz[3], ws.lpart, cl->stream));
    double r, m, p;
    if ((rc = read_loss(ws, cl->stream, &r, &m, &p))) return rc;
    R += r;
    M += m;
    Pp += p;
  }
  if (mse_plain) *mse_plain = M > 0 ? Pp / M : 0.0;
  if (mse_learned) *mse_learned = M > 0 ? R / M : 0.0;
  return FIP_OK;
}

int fip_closure_attach(fip_closure* cl, fip_sim* coarse, const float* bed_std, float zref, float sat,
                       double mass_limit) {
  if (!cl || !coarse || !bed_std) return fail(FIP_E_INVALID, "null argument");
  CHIP(hipSetDevice(cl->dev));
  fip_closure_detach(coarse);
  auto a = std::make_unique<Attachment>();
  a->sim = coarse;
  a->zref_h = zref;
  a->sat = sat;
  a->limit = mass_limit;
  const fip_grid& g = fip_internal::grid(coarse);
  const size_t spx = (size_t)g.W * g.H;
  a->nparts = cl_correct_parts(g.W, g.H);
  int rc = dalloc(&a->bstd, spx);
  if (!rc) rc = dalloc(&a->o, spx * 2);
  if (!rc) rc = dalloc(&a->part, (size_t)a->nparts);
  if (!rc) rc = dalloc(&a->ms, 1);
  if (!rc) rc = dalloc(&a->mpart, (size_t)2 * a->nparts);
  if (rc) return rc;
  CHIP(hipMemcpy(a->bstd, bed_std, spx * sizeof(float), hipMemcpyHostToDevice));
  CHIP(hipMemset(a->ms, 0, sizeof(ClMass)));
  ClDom hd{g, a->bstd, 0, zref, 0};
  if ((rc = dalloc(&a->dom, 1))) return rc;
  CHIP(hipMemcpy(a->dom, &hd, sizeof hd, hipMemcpyHostToDevice));
  if ((rc = upload_plan({g.W}, {g.H}, cl->dev, &a->items, &a->nitems, 1))) return rc;
  Attachment* ap = a.get();
  cl->att.push_back(std::move(a));
  fip_internal::set_post_hook(coarse, [cl, ap](hipStream_t st, int par) -> int {
    const fip_grid& g = fip_internal::grid(ap->sim);
    // the row march runs the step before's mass check; the tile forms check after the correction
    const bool march = cl->fused_form == 2;
    const ClMassCheck mc{ap->ms, ap->part, ap->mpart, ap->nparts, ap->limit};
    CHIP(cl_fused_launch(ap->dom, 1, g.W, g.H, par, cl->wf, cl->wl1, cl->w, ap->o, st, cl->fused_form, ap->items,
                         ap->nitems, cl->act, march ? &mc : nullptr));
    CHIP(cl_correct_launch(g, par, ap->o, kBound, ap->sat, ap->ms, ap->part, ap->mpart, st));
    if (!march) CHIP(cl_mass_check_launch(g, ap->mpart, ap->ms, ap->limit, ap->part, ap->nparts, st));
    return FIP_OK;
  });
  return FIP_OK;
}

int fip_closure_attach_ensemble(fip_closure* cl, fip_ens* ens, const float* bed_std, const float* zref, float sat,
                                double mass_limit) {
  if (!cl || !ens || !bed_std || !zref) return fail(FIP_E_INVALID, "null argument");
  CHIP(hipSetDevice(cl->dev));
  const int n = fip_internal::ens_count(ens);
  std::vector<int> Ws(n), Hs(n);
  std::vector<long long> off(n + 1, 0);
  for (int s = 0; s < n; s++) {
    const fip_grid& g = fip_internal::grid(fip_internal::ens_member(ens, s));
    Ws[s] = g.W;
    Hs[s] = g.H;
    off[s + 1] = off[s] + (long long)g.W * g.H;
  }
  for (auto it = cl->att.begin(); it != cl->att.end(); ++it)
    if ((*it)->ens == ens) {
      cl->att.erase(it);
      break;
    }
  auto a = std::make_unique<Attachment>();
  a->ens = ens;
  a->sat = sat;
  a->limit = mass_limit;
  a->wmax = *std::max_element(Ws.begin(), Ws.end());
  a->hmax = *std::max_element(Hs.begin(), Hs.end());
  const size_t npx = (size_t)off[n];
  a->nparts = cl_correct_parts(a->wmax, a->hmax);  // per scenario, at the largest size
  int rc = dalloc(&a->bstd, npx);
  if (!rc) rc = dalloc(&a->o, npx * 2);
  if (!rc) rc = dalloc(&a->part, (size_t)a->nparts * n);
  if (!rc) rc = dalloc(&a->zref, (size_t)n);
  if (!rc) rc = dalloc(&a->ms, (size_t)n);
  if (!rc) rc = dalloc(&a->mpart, (size_t)2 * a->nparts * n);
  if (rc) return rc;
  CHIP(hipMemcpy(a->bstd, bed_std, npx * sizeof(float), hipMemcpyHostToDevice));
  CHIP(hipMemcpy(a->zref, zref, n * sizeof(float), hipMemcpyHostToDevice));
  CHIP(hipMemset(a->ms, 0, n * sizeof(ClMass)));
  std::vector<ClDom> hd(n);
  for (int s2 = 0; s2 < n; s2++)
    hd[s2] = ClDom{fip_internal::grid(fip_internal::ens_member(ens, s2)), a->bstd + off[s2], off[s2], zref[s2], 0};
  if ((rc = dalloc(&a->dom, (size_t)n))) return rc;
  CHIP(hipMemcpy(a->dom, hd.data(), n * sizeof(ClDom), hipMemcpyHostToDevice));
  if ((rc = upload_plan(Ws, Hs, cl->dev, &a->items, &a->nitems, n))) return rc;
  Attachment* ap = a.get();
  cl->att.push_back(std::move(a));
  fip_internal::ens_set_post_hook(ens, [cl, ap, n](hipStream_t st, int par) -> int {
    const fip_ens_scn* scn = fip_internal::ens_scn(ap->ens);
    const bool march = cl->fused_form == 2;
    const ClMassCheck mc{ap->ms, ap->part, ap->mpart, ap->nparts, ap->limit};
    CHIP(cl_fused_launch(ap->dom, n, ap->wmax, ap->hmax, par, cl->wf, cl->wl1, cl->w, ap->o, st, cl->fused_form,
                         ap->items, ap->nitems, cl->act, march ? &mc : nullptr));
    CHIP(cl_correct_ens_launch(scn, ap->dom, par, ap->o, kBound, ap->sat, ap->ms, ap->part, ap->mpart, ap->wmax,
                               ap->hmax, n, st));
    if (!march) CHIP(cl_mass_check_ens_launch(scn, n, ap->mpart, ap->ms, ap->limit, ap->part, ap->nparts, st));
    return FIP_OK;
  });
  return FIP_OK;
}

int fip_closure_detach(fip_sim* coarse) {
  if (!coarse) return fail(FIP_E_INVALID, "null argument");
  fip_internal::set_post_hook(coarse, nullptr);
  for (fip_closure* cl : registry())
    for (auto it = cl->att.begin(); it != cl->att.end(); ++it)
      if ((*it)->sim == coarse) {
        cl->att.erase(it);
        break;
      }
  return FIP_OK;
}

int fip_closure_status(fip_closure* cl, fip_sim* coarse, fip_ens* ens, int s, int* fallback, unsigned* nsat,
                       double* max_rel) {
  if (!cl) return fail(FIP_E_INVALID, "null argument");
  Attachment* a = find_att(cl, coarse, ens);
  if (!a) return fail(FIP_E_STATE, "not attached");
  if (ens && (s < 0 || s >= fip_internal::ens_count(ens))) return fail(FIP_E_INVALID, "scenario out of range");
  CHIP(hipSetDevice(cl->dev));
  CHIP(hipDeviceSynchronize());
  // the check of the last corrected step, if the next step's network kernel has not run it
  if (ens)
    CHIP(cl_mass_check_ens_launch(fip_internal::ens_scn(ens), fip_internal::ens_count(ens), a->mpart, a->ms, a->limit,
                                  a->part, a->nparts, nullptr));
  else
    CHIP(cl_mass_check_launch(fip_internal::grid(coarse), a->mpart, a->ms, a->limit, a->part, a->nparts, nullptr));
  CHIP(hipDeviceSynchronize());
  ClMass m;
  CHIP(hipMemcpy(&m, a->ms + (ens ? s : 0), sizeof m, hipMemcpyDeviceToHost));
  if (fallback) *fallback = m.fallback;
  if (nsat) *nsat = m.nsat;
  if (max_rel) *max_rel = m.max_rel;
  return FIP_OK;
}

int fip_closure_forward(fip_closure* cl, const float* feat, int N, int H, int W, float* out) {
  if (!cl || !feat || !out) return fail(FIP_E_INVALID, "null argument");
  CHIP(hipSetDevice(cl->dev));
  Work& ws = cl->host;
  int rc = ws.ensure(N, H, W, true);
  if (rc) return rc;
  const long npix = ws.pixels();
  float* df = nullptr;
  CHIP(hipMalloc(&df, npix * CL_FEAT * sizeof(float)));
  CHIP(hipMemcpy(df, feat, npix * CL_FEAT * sizeof(float), hipMemcpyHostToDevice));
  hipError_t e = cl_feat_from_float_launch(df, npix, ws.x0, cl->stream);
  if (e == hipSuccess) rc = forward(cl, ws, nullptr, 0, cl->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(cl->stream);
  hipFree(df);
  if (e != hipSuccess) return fail(FIP_E_HIP, std::string("forward: ") + hipGetErrorString(e));
  if (rc) return rc;
  CHIP(hipMemcpy(out, ws.o, npix * 2 * sizeof(float), hipMemcpyDeviceToHost));
  return FIP_OK;
}

int fip_closure_forward_noalloc(fip_closure* cl, const float* feat, int N, int H, int W, float* out) {
  if (!cl || !feat || !out) return fail(FIP_E_INVALID, "null argument");
  if (N < 1 || H < 1 || W < 1) return fail(FIP_E_INVALID, "empty image batch");
  if (N > 65535) return fail(FIP_E_INVALID, "at most 65535 images per call");
  const long npix = (long)N * H * W;
  if (npix > FIP_CL_NOALLOC_MAX_PIXELS)
    return fail(FIP_E_INVALID, "N*H*W = " + std::to_string(npix) + " pixels exceeds the scratch of " +
                                   std::to_string(FIP_CL_NOALLOC_MAX_PIXELS) + " pixels");
  CHIP(hipSetDevice(cl->dev));
  u16* base = (u16*)g_scratch_ptr[cl->dev];
  if (!base) return fail(FIP_E_STATE, "scratch not set up on this device");
  u16* x0 = base;
  u16* z = base + kNoAllocImage;
  u16* a[2] = {base + 2 * kNoAllocImage, base + 3 * kNoAllocImage};
  CHIP(cl_feat_from_float_launch(feat, npix, x0, cl->stream));
  const u16* in = x0;
  for (int l = 0; l < 3; l++) {
    CHIP(cl_launch_conv(EPI_ACT, cl->act, in, cl->wf + (size_t)l * CL_PACKED, bias_of(cl, l), nullptr, z,
                        a[l & 1], nullptr, N, H, W, nullptr, 0, cl->stream));
    in = a[l & 1];
  }
  CHIP(cl_launch_conv(EPI_OUT, cl->act, in, cl->wf + 3 * (size_t)CL_PACKED, bias_of(cl, 3), nullptr, nullptr,
                      nullptr, out, N, H, W, nullptr, 0, cl->stream));
  CHIP(hipStreamSynchronize(cl->stream));
  return FIP_OK;
}

int fip_closure_backward(fip_closure* cl, const float* feat, int N, int H, int W, const float* qp, const float* qt,
                         const unsigned char* mask, float bound, float* grad, double* loss) {
  if (!cl || !feat || !qp || !qt || !mask || !grad) return fail(FIP_E_INVALID, "null argument");
  CHIP(hipSetDevice(cl->dev));
  Work& ws = cl->host;
  int rc = ws.ensure(N, H, W, true);
  if (rc) return rc;
  const long npix = ws.pixels();
  float* df = nullptr;
  CHIP(hipMalloc(&df, npix * CL_FEAT * sizeof(float)));
  CHIP(hipMemcpy(df, feat, npix * CL_FEAT * sizeof(float), hipMemcpyHostToDevice));
  CHIP(hipMemcpy(ws.qp, qp, npix * 2 * sizeof(float), hipMemcpyHostToDevice));
  CHIP(hipMemcpy(ws.qt, qt, npix * 2 * sizeof(float), hipMemcpyHostToDevice));
  CHIP(hipMemcpy(ws.mask, mask, npix * 2, hipMemcpyHostToDevice));
  hipError_t e = cl_feat_from_float_launch(df, npix, ws.x0, cl->stream);
  if (e == hipSuccess) rc = forward(cl, ws, nullptr, 0, cl->stream);
  if (e == hipSuccess && !rc) rc = backward(cl, ws, bound, cl->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(cl->stream);
  hipFree(df);
  if (e != hipSuccess) return fail(FIP_E_HIP, std::string("backward: ") + hipGetErrorString(e));
  if (rc) return rc;
  double r, m, p;
  if ((rc = read_loss(ws, cl->stream, &r, &m, &p))) return rc;
  if (loss) *loss = r / (2.0 * (double)npix);
  CHIP(hipMemcpy(grad, cl->g, sizeof(float) * CL_LAYERS * CL_LAYER_PARAMS, hipMemcpyDeviceToHost));
  return FIP_OK;
}

int fip_closure_fused_outputs(fip_closure* cl, fip_sim* coarse, const float* bed_std, float zref, float* out) {
  if (!cl || !coarse || !bed_std || !out) return fail(FIP_E_INVALID, "null argument");
  CHIP(hipSetDevice(cl->dev));
  CHIP(hipDeviceSynchronize());
  const fip_grid& g = fip_internal::grid(coarse);
  const size_t spx = (size_t)g.W * g.H;
  float *bstd = nullptr, *o = nullptr;
  ClDom* dom = nullptr;
  CHIP(hipMalloc(&bstd, spx * sizeof(float)));
  CHIP(hipMalloc(&o, spx * 2 * sizeof(float)));
  CHIP(hipMalloc(&dom, sizeof(ClDom)));
  CHIP(hipMemcpy(bstd, bed_std, spx * sizeof(float), hipMemcpyHostToDevice));
  ClDom hd{g, bstd, 0, zref, 0};
  CHIP(hipMemcpy(dom, &hd, sizeof hd, hipMemcpyHostToDevice));
  int4* items = nullptr;
  int nitems = 0;
  if (int rc = upload_plan({g.W}, {g.H}, cl->dev, &items, &nitems)) return rc;
  // par such that par ^ 1 is the current side and the halt flag read is that of no step
  const int par = fip_internal::cur_side(coarse) ^ 1;
  hipError_t e = cl_fused_launch(dom, 1, g.W, g.H, par, cl->wf, cl->wl1, cl->w, o, nullptr, cl->fused_form, items,
                                 nitems, cl->act, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, o, spx * 2 * sizeof(float), hipMemcpyDeviceToHost);
  hipFree(bstd); hipFree(o); hipFree(dom); hipFree(items);
  if (e != hipSuccess) return fail(FIP_E_HIP, std::string("fused outputs: ") + hipGetErrorString(e));
  return FIP_OK;
}

int fip_closure_set_activation(fip_closure* cl, int act) {
  if (!cl) return fail(FIP_E_INVALID, "null argument");
  if (act != FIP_CLOSURE_ACT_GELU && act != FIP_CLOSURE_ACT_HSILU)
    return fail(FIP_E_INVALID, "activation must be 0 (GELU) or 1 (hard SiLU)");
  if (!cl->att.empty()) return fail(FIP_E_INVALID, "set the activation before attaching the closure");
  cl->act = act;
  return FIP_OK;
}

int fip_closure_get_activation(fip_closure* cl, int* act) {
  if (!cl || !act) return fail(FIP_E_INVALID, "null argument");
  *act = cl->act;
  return FIP_OK;
}

int fip_closure_set_kernel(fip_closure* cl, int form) {
  if (!cl) return fail(FIP_E_INVALID, "null argument");
  if (form < 0 || form > 2)
    return fail(FIP_E_INVALID, "kernel form must be 0 (per tile), 1 (persistent tiles) or 2 (row march)");
  // the row march runs a step's mass check in the next step's launch, the tile forms after
  // the correction: a change between two steps would skip a check
  if (!cl->att.empty()) return fail(FIP_E_INVALID, "set the kernel form before attaching the closure");
  cl->fused_form = form;
  return FIP_OK;
}

int fip_mfma_probe(int device, const unsigned short* A, const unsigned short* B, float* D) {
  if (!A || !B || !D) return fail(FIP_E_INVALID, "null argument");
  CHIP(hipSetDevice(device));
  u16 *dA = nullptr, *dB = nullptr;
  float* dD = nullptr;
  CHIP(hipMalloc(&dA, 16 * 32 * 2));
  CHIP(hipMalloc(&dB, 32 * 16 * 2));
  CHIP(hipMalloc(&dD, 16 * 16 * 4));
  CHIP(hipMemcpy(dA, A, 16 * 32 * 2, hipMemcpyHostToDevice));
  CHIP(hipMemcpy(dB, B, 32 * 16 * 2, hipMemcpyHostToDevice));
  CHIP(cl_mfma_probe_launch(dA, dB, dD, nullptr));
  CHIP(hipDeviceSynchronize());
  CHIP(hipMemcpy(D, dD, 16 * 16 * 4, hipMemcpyDeviceToHost));
  hipFree(dA);
  hipFree(dB);
  hipFree(dD);
  return FIP_OK;
}

}  // extern "C"
