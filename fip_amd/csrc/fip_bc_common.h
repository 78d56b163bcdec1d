// Boundary-condition cell bodies, shared by the run-time compiled boundary kernel of one
// domain and the precompiled ensemble kernels (this file is also handed to hipRTC). With
// type, side and staggered known at compile time the switches fold away.
#pragma once
#include "fip_detmath.h"

#define FIP_BCT_Z 0
#define FIP_BCT_Q 1
#define FIP_BCT_CLOSE 2
#define FIP_BCT_OPEN 3
#define FIP_BCS_LEFT 0
#define FIP_BCS_RIGHT 1
#define FIP_BCS_TOP 2
#define FIP_BCS_BOTTOM 3

// One term of BoundaryCondition::sumWaterDepths (src/boundary_condition.cpp:6-18): the
// caller adds the terms of cells from..to-1 in that order and multiplies by dx.
__device__ __forceinline__ float fip_swd_term(float z, float bed) {
  float ah = z - bed;
  if (ah < 0.01f) ah = 0.f;
  return __builtin_sqrtf(9.81f * ah) * ah;
}

// Element index of cell `pos` of a BC side, the cell whose value is copied or
// extrapolated (the neighbour) and the staggered face (src/bc_kernel_gen.cpp:11-73).
__device__ __forceinline__ void fip_bc_index(int side, int pos, int P, int W, int H, long& idx,
                                             long& nb, long& stag) {
  if (side == FIP_BCS_RIGHT) {
    idx = (long)pos * P + (W - 1); stag = idx - 1; nb = idx - 1;
  } else if (side == FIP_BCS_LEFT) {
    idx = (long)pos * P; stag = idx; nb = idx + 1;
  } else if (side == FIP_BCS_TOP) {
    idx = (long)pos + (long)(H - 1) * P; stag = idx - P; nb = idx - P;
  } else {
    idx = (long)pos; stag = idx; nb = idx + P;
  }
}

// The body of one BC at one border position (src/bc_kernel_gen.cpp:75-141): staggered
// bodies for solver rmg, the cell-centred ones for cn. q is qx on the left/right sides
// and qy on top/bottom; q_lateral is the other one.
__device__ __forceinline__ void fip_bc_cell(int type, int side, int staggered, float qval,
                                            float zval, float* h, float* qx, float* qy,
                                            const float* sohle, int P, int W, int H, int pos) {
  long idx, nb, stag;
  fip_bc_index(side, pos, P, W, H, idx, nb, stag);
  const bool xface = side <= FIP_BCS_RIGHT;
  float* q = xface ? qx : qy;
  float* ql = xface ? qy : qx;
  const long qi = staggered ? stag : idx;
  switch (type) {
    case FIP_BCT_Z:
      if (staggered) {
        h[idx] = fip_fmaxf(0.f, zval - sohle[idx]);
        h[nb] = fip_fmaxf(0.f, zval - sohle[nb]);
        q[stag] = -0.f;
      } else {
        const float qnb = q[nb], hnb = h[nb];
        const float ah = fip_fmaxf(0.f, zval - sohle[idx]);
        h[idx] = ah;
        q[idx] = 0.99f * (qnb / fip_fmaxf(hnb, 0.01f)) * ah;
      }
      break;
    case FIP_BCT_Q: {
      const float bidx = sohle[idx];
      const float hnb = h[nb], bnb = sohle[nb];
      float ah = zval - bidx;
      if (ah < 0.f) ah = 0.f;
      q[qi] = __builtin_sqrtf(9.81f * ah) * ah * qval;
      if (ah > 0.f) {
        if (staggered) {
          const float z_neighbor = hnb + bnb;
          h[idx] = fip_fmaxf(ah / 2, z_neighbor - bidx);
        } else {
          h[idx] = fip_fmaxf(ah / 2, hnb + bnb - bidx);
        }
      }
      break;
    }
    case FIP_BCT_CLOSE: {
      const float qlnb = ql[nb], hnb = h[nb];
      q[qi] = 0.f;
      ql[idx] = qlnb;
      h[idx] = hnb;
      break;
    }
    default: {  // open: the discharge points out of the domain, +1 right/top, -1 left/bottom
      const float sign = (side == FIP_BCS_RIGHT || side == FIP_BCS_TOP) ? 1.0f : -1.0f;
      const float ah_neighbor = h[nb];
      const float z_neighbor = ah_neighbor + sohle[nb];
      if (ah_neighbor > 0.f) {
        q[qi] = sign * __builtin_sqrtf(9.81f * ah_neighbor) * ah_neighbor;
        h[idx] = fip_fminf(ah_neighbor, fip_fmaxf(0.f, z_neighbor - sohle[idx]));
      }
      break;
    }
  }
}
