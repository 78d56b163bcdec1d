// Ensemble kernels around the ensemble stencil: the per-scenario post-step control and
// BC parameters, the BC application over all scenarios, the global active-tile schedule
// and the flood plains. Built with -ffp-contract=off: every scenario computes the same
// bits as a single-domain run of it (the same cell bodies, series interpolation,
// sumWaterDepths order and dt control as the run-time compiled boundary kernel).
#include <hip/hip_runtime.h>

#include "fip_bc_common.h"
#include "fip_ctl.h"
#include "fip_detmath.h"
#include "fip_sched.h"
#include "fip_step_common.h"

namespace {

// getZ / getQ of one BC at time tf (src/boundary_condition.cpp:45-77): a BC with a
// series reads z (and q for type q) from it, 0 where that series is missing.
__device__ __forceinline__ bool ens_bc_series(const fip_ens_bc& b) { return b.zn > 0 || b.qn > 0; }

__device__ __forceinline__ float ens_bc_z(const fip_ens_bc& b, const float* ser, float tf) {
  if (!ens_bc_series(b)) return b.z;
  return b.zn > 0 ? fip_interp(ser + b.zoff, ser + b.zoff + b.zn, b.zn, tf) : 0.f;
}

__device__ __forceinline__ float ens_bc_qnum(const fip_ens_bc& b, const float* ser, float tf) {
  if (!ens_bc_series(b)) return b.q;
  return b.qn > 0 ? fip_interp(ser + b.qoff, ser + b.qoff + b.qn, b.qn, tf) : 0.f;
}

}  // namespace

// Post-step control of every scenario, one work-group each (modes as the single-domain
// boundary kernel: 0 = before the first step, 1 = after a step of parity par). Unlike a
// single domain, an ensemble has no sampling or save_state; a scenario stops when its
// duration reaches its target.
extern "C" __global__ __launch_bounds__(256) void fip_ens_post(const fip_ens_scn* scn,
                                                               const fip_ens_bc* bcs,
                                                               const float* series,
                                                               fip_ens_ctl* ec, int mode,
                                                               int par) {
  const int sidx = blockIdx.x;
  const int tid = threadIdx.x;
  const fip_grid g = scn[sidx].g;
  const int bc0 = scn[sidx].bc0, nbc = scn[sidx].nbc;
  fip_ctl* c = g.ctl;
  __shared__ int s_skip;
  __shared__ float s_zv[FIP_MAX_BC], s_swd[FIP_MAX_BC];
  __shared__ int s_need[FIP_MAX_BC];
  __shared__ float s_terms[256];
  __shared__ fip_ens_bc s_bc[FIP_MAX_BC];
  // the scenario's BC descriptors, staged by the whole work-group: thread 0's serial part
  // then reads them from LDS instead of one dependent global load per BC
  for (int i = tid; i < nbc; i += blockDim.x) s_bc[i] = bcs[bc0 + i];
  __syncthreads();
  if (sidx == 0 && tid == 0 && mode == 1) {
    ec->list_count[par] = 0u;
    ec->work_next[par] = 0u;
  }
  double dur = 0.0, target = 0.0;
  float dt = 0.f, cflp = 0.f;
  unsigned acc = 0u;
  int it = 0, vdt = 0, hreq = 0;
  if (tid == 0) {
    hreq = c->halt_req[par ^ 1]; dur = c->duration; dt = c->dt; it = c->iteration;
    acc = c->cfl_acc[par]; cflp = c->cfl_prev; vdt = c->variable_dt; target = c->target;
  }
  if (tid == 0) {
    int skip = 0;
    if (mode == 1) {
      if (hreq) {
        c->halt_req[par] = hreq;
        skip = 1;
      } else {
        dur = dur + (double)dt;
        it = it + 1;
        cflp = fip_u2f(acc);
        c->duration = dur; c->iteration = it; c->cfl_prev = cflp;
        c->cfl_acc[par] = fip_f2u(1.0f); c->list_count[par] = 0u; c->work_next[par] = 0u;
      }
    } else {
      if (vdt) { dt = fip_control_dt(dt, cflp); c->dt = dt; }
    }
    if (!skip) {
      const float tf = (float)dur;
      for (int b = 0; b < nbc; b++) {
        const fip_ens_bc& B = s_bc[b];
        const float zv = ens_bc_z(B, series, tf);
        s_zv[b] = zv;
        s_need[b] = B.type == FIP_BCT_Q &&
                    !(c->swd_valid[b] && fip_f2u(c->swd_z[b]) == fip_f2u(zv));
      }
    }
    s_skip = skip;
  }
  __syncthreads();
  if (s_skip) return;
  // sumWaterDepths where z changed: terms in parallel, summed by thread 0 in order
  for (int b = 0; b < nbc; b++) {
    if (!s_need[b]) continue;
    const fip_ens_bc& B = s_bc[b];
    const float zv = s_zv[b];
    float ah_sum = 0.f;
    for (int base = B.from; base < B.to; base += 256) {
      const int k = base + tid;
      if (k < B.to) {
        long idx, nb, stag;
        fip_bc_index(B.side, k, g.P, g.W, g.H, idx, nb, stag);
        s_terms[tid] = fip_swd_term(zv, g.bed[idx]);
      }
      __syncthreads();
      if (tid == 0) {
        const int m = B.to - base < 256 ? B.to - base : 256;
        for (int e = 0; e < m; e++) ah_sum += s_terms[e];
      }
      __syncthreads();
    }
    if (tid == 0) s_swd[b] = ah_sum * g.dx;
  }
  if (tid != 0) return;
  const float tf = (float)dur;
  for (int b = 0; b < nbc; b++) {
    const fip_ens_bc& B = s_bc[b];
    float qv;
    if (B.type == FIP_BCT_Q) {
      float swd = c->swd_val[b];
      if (s_need[b]) {
        swd = s_swd[b];
        c->swd_z[b] = s_zv[b]; c->swd_val[b] = swd; c->swd_valid[b] = 1;
      }
      qv = ens_bc_qnum(B, series, tf) / swd;
    } else {
      qv = ens_bc_series(B) ? 0.f : B.q;
    }
    c->bc_q[b] = qv;
    c->bc_z[b] = s_zv[b];
  }
  if (mode == 1) {
    const int req = (dur < target) ? 0 : FIP_HALT_FINISHED;
    c->halt_req[par] = req;
    ec->finished[sidx] = req ? 1 : 0;
    if (!req && vdt) c->dt = fip_control_dt(dt, cflp);
  }
}

// Global active-tile schedule of list q: every tile of every running scenario whose
// dilated mask is non-empty (all valid rows in dense mode), one atomic per work-group and
// pass (fip_block_append; every thread of the work-group runs the same passes).
__device__ __forceinline__ void ens_schedule(const fip_ens_scn* scn, int nscn, int total,
                                            fip_ens_ctl* ec, unsigned long long* list, int q,
                                            int dense, int block, int nblocks) {
  const int stride = nblocks * blockDim.x;
  for (int base = block * blockDim.x; base < total; base += stride) {
    const int gt = base + threadIdx.x;
    unsigned long long need = 0ull;
    int s = 0, t = 0;
    if (gt < total) {
      int lo = 0, hi = nscn - 1;  // last scenario with tile_base <= gt
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (scn[mid].tile_base <= gt) lo = mid;
        else hi = mid - 1;
      }
      s = lo;
      const fip_grid& g = scn[s].g;
      t = gt - scn[s].tile_base;
      if (!g.ctl->halt_req[q ^ 1]) {
        if (dense) {
          const int R = g.tile_rows, ty = t / g.tiles_x;
          const int nrows = (g.H - ty * R) < R ? (g.H - ty * R) : R;
          need = fip_valid_rows(nrows);
        } else {
          need = fip_tile_need(g, t);
        }
      }
    }
    const bool act = need != 0ull;
    const unsigned pos = fip_block_append(act, &ec->list_count[q]);
    if (act && pos < (unsigned)total) {
      const unsigned long long lo = (unsigned long long)__builtin_ctzll(need);
      const unsigned long long hi = 63ull - (unsigned long long)__builtin_clzll(need);
      list[pos] = ((unsigned long long)s << 40) | ((unsigned long long)t << 12) | (lo << 6) | hi;
    }
  }
}

// BC application of one stage for all scenarios: work-group b < nchunks applies chunk b
// (up to 256 positions of one BC). Work-groups beyond nchunks build list `q` when
// sched_q >= 0.
extern "C" __global__ __launch_bounds__(256) void fip_ens_apply(
    const fip_ens_scn* scn, int nscn, int total_tiles, const fip_ens_bc* bcs,
    const fip_ens_chunk* chunks, int nchunks, fip_ens_ctl* ec, unsigned long long* list,
    int mode, int par, int sched_q, int dense) {
  if ((int)blockIdx.x >= nchunks) {
    if (sched_q >= 0)
      ens_schedule(scn, nscn, total_tiles, ec, list, sched_q, dense, blockIdx.x - nchunks,
                   gridDim.x - nchunks);
    return;
  }
  const fip_ens_chunk ch = chunks[blockIdx.x];
  if ((int)threadIdx.x >= ch.n) return;
  const fip_grid& g = scn[ch.scn].g;
  const fip_ctl* c = g.ctl;
  if (mode == 1 && c->halt_req[par ^ 1]) return;
  const int cur = mode == 1 ? (par ^ 1) : par;
  const fip_ens_bc B = bcs[ch.bc];
  const int b = ch.bc - scn[ch.scn].bc0;
  const float qval = c->bc_q[b], zval = c->bc_z[b];
  fip_bc_cell(B.type, B.side, scn[ch.scn].staggered, qval, zval, g.h[cur], g.qx[cur], g.qy[cur],
              g.bed, g.P, g.W, g.H, ch.pos0 + (int)threadIdx.x);
}

// flood_plains_kernel (src/kernels.cu:254-262) of every scenario whose iteration is a
// multiple of 100 (src/main.cpp:125-128); idempotent. blockIdx.y = scenario.
extern "C" __global__ void fip_ens_floodplain(const fip_ens_scn* scn) {
  const fip_grid& g = scn[blockIdx.y].g;
  const int it = g.ctl->iteration;
  if ((it % 100) != 0) return;
  const float* h = g.h[it & 1];
  const long n = (long)g.W * g.H;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    const long i = (k / g.W) * g.P + (k % g.W);
    g.hmax[i] = fip_fmaxf(g.hmax[i], h[i]);
  }
}

hipError_t fip_launch_ens(int which, const fip_ens_scn* scn, int nscn, int total_tiles,
                          const fip_ens_bc* bcs, const float* series, const fip_ens_chunk* chunks,
                          int nchunks, fip_ens_ctl* ec, unsigned long long* list, int mode, int par,
                          int sched_q, int dense, int blocks, hipStream_t st) {
  switch (which) {
    case 0:
      hipLaunchKernelGGL(fip_ens_post, dim3(nscn), dim3(256), 0, st, scn, bcs, series, ec, mode, par);
      break;
    case 1:
      hipLaunchKernelGGL(fip_ens_apply, dim3(blocks), dim3(256), 0, st, scn, nscn, total_tiles, bcs,
                         chunks, nchunks, ec, list, mode, par, sched_q, dense);
      break;
    case 2:
      hipLaunchKernelGGL(fip_ens_floodplain, dim3(blocks, nscn), dim3(256), 0, st, scn);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
