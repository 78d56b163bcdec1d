// Small kernels around the stencil: the fine-grained step bookkeeping, controlTimestep,
// the flood-plain maximum and the active-tile scheduler.
#include <hip/hip_runtime.h>

#include "fip_ctl.h"
#include "fip_detmath.h"
#include "fip_sched.h"
#include "fip_step_common.h"

extern "C" __global__ void fip_k_control(fip_ctl* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) fip_control(c);
}

// End of one fine-grained stencil step of parity par: duration += dt (kernels.cu:454),
// iteration++ and the ping-pong swap (main.cpp:118-121), and the CFL minimum of the step.
extern "C" __global__ void fip_k_advance(fip_grid g, int par) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  fip_ctl* c = g.ctl;
  c->duration = c->duration + (double)c->dt;
  c->iteration = c->iteration + 1;
  c->cfl_prev = fip_u2f(c->cfl_acc[par]);
  c->cfl_acc[par] = fip_f2u(1.0f);
  c->list_count[par] = 0u;
  c->work_next[par] = 0u;
}

extern "C" __global__ void fip_k_schedule(fip_grid g, int q) {
  fip_schedule(g, q, blockIdx.x, gridDim.x);
}

// flood_plains_kernel (src/kernels.cu:254-262) on the current buffers. With check set it
// runs only when iteration % 100 == 0 (src/main.cpp:125-128); it is idempotent.
extern "C" __global__ void fip_k_floodplain(fip_grid g, int check) {
  const fip_ctl* c = g.ctl;
  const int it = c->iteration;
  if (check && (it % 100) != 0) return;
  const float* h = g.h[it & 1];
  const int n = g.W;
  for (int row = blockIdx.y; row < g.H; row += gridDim.y)
    for (int col = blockIdx.x * blockDim.x + threadIdx.x; col < n; col += gridDim.x * blockDim.x) {
      const long i = (long)row * g.P + col;
      g.hmax[i] = fip_fmaxf(g.hmax[i], h[i]);
    }
}

// Self-check of fip_rcp_rn: every binary32 b with bits in [lo, lo + n) against the IEEE
// division 1.0f / b; counts the differences and keeps the smallest such bit pattern.
extern "C" __global__ void fip_k_check_rcp(unsigned lo, unsigned n, unsigned long long* nbad, unsigned* first) {
  unsigned long long bad = 0;
  unsigned lowest = 0xffffffffu;
  for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const float b = fip_u2f(lo + k);
    if (fip_f2u(fip_rcp_rn(b)) != fip_f2u(1.0f / b)) {
      bad++;
      lowest = lowest < lo + k ? lowest : lo + k;
    }
  }
  if (bad) {
    atomicAdd(nbad, bad);
    atomicMin(first, lowest);
  }
}

hipError_t fip_launch_check_rcp(unsigned lo, unsigned n, unsigned long long* nbad, unsigned* first, hipStream_t st) {
  hipLaunchKernelGGL(fip_k_check_rcp, dim3(4096), dim3(256), 0, st, lo, n, nbad, first);
  return hipGetLastError();
}

hipError_t fip_launch_misc(int which, const fip_grid& g, int arg, int blocks, hipStream_t st) {
  switch (which) {
    case 0: hipLaunchKernelGGL(fip_k_control, dim3(1), dim3(64), 0, st, g.ctl); break;
    case 1: hipLaunchKernelGGL(fip_k_advance, dim3(1), dim3(64), 0, st, g, arg); break;
    case 2: hipLaunchKernelGGL(fip_k_schedule, dim3(blocks), dim3(256), 0, st, g, arg); break;
    case 3: {
      const int gx = (g.W + 255) / 256 < 4 ? (g.W + 255) / 256 : 4;
      const int gy = g.H < 1024 ? g.H : 1024;
      hipLaunchKernelGGL(fip_k_floodplain, dim3(gx, gy), dim3(256), 0, st, g, arg);
      break;
    }
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
