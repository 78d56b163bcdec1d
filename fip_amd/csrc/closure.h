// Shared by the closure kernels (closure.hip) and their host code (closure_host.cpp).
#pragma once
#include <vector>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fip_ctl.h"

#define CL_C 32          // channels of every activation tensor
#define CL_TAPS 9
#define CL_FEAT 6        // real input features
#define CL_OUT 2         // real outputs (corrections of qx and qy)
#define CL_LAYERS 4
#define CL_LAYER_PARAMS (CL_TAPS * CL_C * CL_C + CL_C)
#define CL_PACKED (2 * CL_TAPS * 64 * 8)   // bf16 elements of one packed layer

typedef unsigned short u16;
enum { EPI_ACT = 0, EPI_OUT = 1, EPI_DACT = 2 };
// the network's activation (include/fip_closure.h: FIP_CLOSURE_ACT_*)
enum { CL_ACT_GELU = 0, CL_ACT_HSILU = 1 };

// Input features of one coarse state: z - zref, qx, qy, bed - zref, subgrid bed std, kSt.
struct ClFeat {
  const float* h;
  const float* qx;
  const float* qy;
  const float* bed;
  const float* kst;    // NULL = uniform
  const float* bstd;   // W x H dense
  float kst_u, zref;
  int W, H, P;
};

struct ClMass {
  double v_prev;     // volume of the strict interior after the previous step
  double t_prev;     // duration after the previous step
  double max_rel;    // largest relative mass error seen
  double last_rel;
  int valid;
  int fallback;      // set once the relative mass error exceeded the limit
  unsigned nsat;     // faces that kept the plain step because the network saturated
  unsigned steps;
  double t_pend;     // duration after the corrected step whose check is pending
  int pending;       // the correction wrote its slots; the check has not run yet
  int pad;
};


// The pending mass checks that the row-march network kernel runs (one extra work-group
// per domain): each domain's nparts slots in part and mpart, its ClMass in ms.
struct ClMassCheck {
  ClMass* ms;
  const unsigned* part;
  const double* mpart;
  int nparts;
  double limit;
};

// One coarse domain of the fused inference step.
struct ClDom {
  fip_grid g;
  const float* bstd;   // W x H dense
  long long o_off;     // pixel offset of the domain's output image ([H][W][2]) in o
  float zref;
  int pad;
};
#define CL_PACKED_L1 (2 * 3 * 64 * 8)   // bf16 elements of the 8-channel packing of layer 0

// Network outputs o of every cell of ndom domains (domain d's [H][W][2] image at pixel
// offset dom[d].o_off; the domains may differ in size, W x H bounds them all). form 2:
// row march over the work items of cl_fused_plan (items, nitems), 1: persistent tiles,
// 0: one work-group per tile. act: CL_ACT_GELU or CL_ACT_HSILU. mc (form 2, or null): the
// pending mass checks of the step before, run by one more work-group per domain.
hipError_t cl_fused_launch(const ClDom* dom, int ndom, int W, int H, int par, const u16* wf, const u16* wl1,
                           const float* w, float* o, hipStream_t st, int form, const int4* items, int nitems,
                           int act, const ClMassCheck* mc);
// The row march's work items (domain, first column, first row, end row) for domains of
// sizes W[d] x H[d] on a device of `cus` compute units: strips of the kernel's width and
// spans of one length, chosen so that the items (one work-group per compute unit at a
// time) finish in the fewest march steps.
std::vector<int4> cl_fused_plan(const int* W, const int* H, int ndom, int cus);
hipError_t cl_pack_l1_launch(const float* w, u16* wl1, hipStream_t st);
hipError_t cl_launch_conv(int epi, int act, const u16* X, const u16* wp, const float* bias, const u16* Zp, u16* Z,
                          u16* A, float* out, int N, int H, int W, const fip_ctl* ctl, int par,
                          hipStream_t st);
hipError_t cl_mfma_probe_launch(const u16* A, const u16* B, float* D, hipStream_t st);
hipError_t cl_restrict_launch(fip_grid f, int fcur, fip_grid c, int ccur, int F, hipStream_t st);
hipError_t cl_restrict_flux_launch(fip_grid f, int fcur, int F, int Wc, int Hc, float* qt, uint8_t* mask,
                                   hipStream_t st);
hipError_t cl_flux_launch(fip_grid g, int cur, float* q, hipStream_t st);
hipError_t cl_features_launch(ClFeat s, u16* out, int cstride, const fip_ctl* ctl, int par, hipStream_t st);
hipError_t cl_gather_launch(const u16* samples, long sample_elems, int W, const int* origin, int ps, int ph,
                            u16* out, int nb, hipStream_t st);
hipError_t cl_gather_targets_launch(const float* qp, const float* qt, const uint8_t* mask, long sample_pixels,
                                    int W, const int* origin, int ps, int ph, float* bqp, float* bqt,
                                    uint8_t* bmask, int nb, hipStream_t st);
hipError_t cl_feat_from_float_launch(const float* f, long npix, u16* out, hipStream_t st);
hipError_t cl_wgrad_launch(const u16* X, const u16* dZ, float* part, int N, int H, int W, int nblocks,
                           hipStream_t st);
hipError_t cl_reduce_launch(const float* part, int nparts, int n, float* out, hipStream_t st);
hipError_t cl_loss_launch(const float* o, const float* qp, const float* qt, const uint8_t* mask, long npix,
                          float bound, float inv_count, u16* dz, double* part, hipStream_t st);
hipError_t cl_pack_launch(const float* w, int nlayers, u16* wf, u16* wb, hipStream_t st);
hipError_t cl_loss_sum_launch(const double* part, int nb, float* out, hipStream_t st);
hipError_t cl_adam_launch(float* w, float* m, float* v, const float* g, int nlayers, float lr, float b1,
                          float b2, float eps, float c1, float c2, hipStream_t st);
// Correction of the current face discharges by o; saturated faces counted and the
// mass-balance terms of the uncorrected state summed per work-group of 256 columns of
// CL_CORRECT_ROWS rows into part and mpart (2 doubles per slot; cl_correct_parts(W, H)
// slots per domain), which the mass check adds in a fixed order.
#define CL_CORRECT_ROWS 8
inline int cl_correct_parts(int W, int H) { return (W + 255) / 256 * ((H + CL_CORRECT_ROWS - 1) / CL_CORRECT_ROWS); }
hipError_t cl_correct_launch(fip_grid g, int par, const float* o, float bound, float sat, ClMass* ms,
                             unsigned* part, double* mpart, hipStream_t st);
hipError_t cl_mass_check_launch(fip_grid g, const double* mpart, ClMass* ms, double limit, const unsigned* part,
                                int nparts, hipStream_t st);
// (ensembles: dom[s].o_off places scenario s's outputs in o; W x H bounds every scenario,
// and a scenario's work-groups past its own size write empty slots)
hipError_t cl_correct_ens_launch(const fip_ens_scn* scn, const ClDom* dom, int par, const float* o, float bound,
                                 float sat, ClMass* ms, unsigned* part, double* mpart, int W, int H, int n,
                                 hipStream_t st);
hipError_t cl_mass_check_ens_launch(const fip_ens_scn* scn, int n, const double* mpart, ClMass* ms, double limit,
                                    const unsigned* part, int nparts, hipStream_t st);
