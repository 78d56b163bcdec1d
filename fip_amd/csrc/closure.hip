// Learned subgrid closure on gfx950: 4x restriction of a fine domain onto a coarse one,
// the CNN that corrects coarse face discharges (4 layers of 3x3 convolutions, 32
// channels, GELU or hard SiLU, bounded tanh output), its backward pass and Adam.
//
// Convolutions are implicit GEMMs on the bf16 matrix cores (v_mfma_f32_16x16x32_bf16,
// fp32 accumulation). Activations are NHWC bf16 with 32 channels (the 6 input features
// and the 2 outputs are zero-padded to 32, so every layer is one kernel shape). For
// D[co][pixel] = sum_k A[co][k] B[k][pixel] with k = (tap, ci):
//   A fragment: lane l holds A[l % 16][8 (l / 16) + j], j = 0..7 (packed weights)
//   B fragment: lane l holds B[8 (l / 16) + j][l % 16]    (8 channels of one pixel)
//   D:          lane l holds D[4 (l / 16) + r][l % 16], r = 0..3 (4 channels of a pixel)
// The closure has no counterpart in the reference.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <type_traits>

#include "closure.h"
#include "fip_ctl.h"
#include "fip_detmath.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

// float -> bf16, round to nearest even (the native conversion of gfx950)
__device__ __forceinline__ u16 f2bf(float f) { return __builtin_bit_cast(u16, (__bf16)f); }
__device__ __forceinline__ float bf2f(u16 b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }
// two floats -> bf16 pair (a in the low half), one conversion instruction
__device__ __forceinline__ uint32_t pack_bf2(float a, float b) {
  typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, (bf2){(__bf16)a, (__bf16)b});
}

// tanh(u) = 1 - 2 / (1 + e^(2u)) with the hardware exp and reciprocal (saturates to +-1)
__device__ __forceinline__ float fast_tanh(float u) {
  return 1.f - 2.f * __builtin_amdgcn_rcpf(1.f + __expf(2.f * u));
}

// GELU, tanh form: 0.5 x (1 + tanh(s (x + 0.044715 x^3))), s = sqrt(2/pi), evaluated as
// x / (1 + 2^(k x (1 + 0.044715 x^2))) with k = -2 s log2(e) (0.5 (1 + tanh u) is the
// logistic function of 2u): seven instructions, two of them transcendental. Saturates to
// x and to -0 for large |x|.
__device__ __forceinline__ float gelu(float x) {
  constexpr float k = -2.f * 0.7978845608028654f * 1.4426950408889634f;
  const float t = x * __builtin_fmaf(k * 0.044715f, x * x, k);
  return x * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(t));
}
__device__ __forceinline__ float dgelu(float x) {
  const float s = 0.7978845608028654f;
  const float u = s * (x + 0.044715f * x * x * x);
  const float t = fast_tanh(u);
  return 0.5f * (1.f + t) + 0.5f * x * (1.f - t * t) * s * (1.f + 3.f * 0.044715f * x * x);
}

// Hard SiLU: x * clamp(x / 4 + 1 / 2, 0, 1), x times the logistic function's tangent at 0
// clamped to [0, 1]: one fused multiply-add with the output clamp and one multiply, no
// transcendental. x / 4 is exact, so the value is the binary32 rounding of
// x * RN(x / 4 + 1 / 2) on any IEEE machine. Derivative: 0 for x <= -2, x / 2 + 1 / 2
// between, 1 for x >= 2.
__device__ __forceinline__ float hsilu(float x) {
  return x * fminf(fmaxf(__builtin_fmaf(x, 0.25f, 0.5f), 0.f), 1.f);
}
__device__ __forceinline__ float dhsilu(float x) {
  return x >= 2.f ? 1.f : x <= -2.f ? 0.f : __builtin_fmaf(x, 0.5f, 0.5f);
}

// the network's activation (CL_ACT_GELU or CL_ACT_HSILU) and its derivative
template <int ACT>
__device__ __forceinline__ float act(float x) {
  if constexpr (ACT == CL_ACT_HSILU)
    return hsilu(x);
  else
    return gelu(x);
}
template <int ACT>
__device__ __forceinline__ float dact(float x) {
  if constexpr (ACT == CL_ACT_HSILU)
    return dhsilu(x);
  else
    return dgelu(x);
}

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 ld8(const u16* p) {  // 16-byte aligned
  const uint4 v = *(const uint4*)p;
  return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0)); }

}  // namespace

// ---------------------------------------------------------------------------------------
// Layout check of the matrix instruction: D = A B for 16x16x32 operands given row-major
// in A (16 x 32) and B (32 x 16); one wavefront. Used by a test before the convolutions.
extern "C" __global__ void cl_mfma_probe(const u16* A, const u16* B, float* D) {
  const int l = threadIdx.x;
  __attribute__((aligned(16))) u16 a[8], b[8];
  for (int j = 0; j < 8; j++) {
    a[j] = A[(l % 16) * 32 + 8 * (l / 16) + j];
    b[j] = B[(8 * (l / 16) + j) * 16 + (l % 16)];
  }
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = mfma(ld8(a), ld8(b), acc);
  for (int r = 0; r < 4; r++) D[(4 * (l / 16) + r) * 16 + (l % 16)] = acc[r];
}

// ---------------------------------------------------------------------------------------
// Restriction by a factor F of a fine domain's current state onto a coarse domain (both
// pitched fip_grid layouts): h = block mean of h (so the coarse volume is the fine one);
// qx on the coarse east face = mean of the F fine east faces it covers (total discharge
// through the face is kept), qy likewise on the north face. Both ping-pong sides of the
// coarse domain receive the state; the coarse duration becomes the fine one.
extern "C" __global__ void cl_restrict(fip_grid f, int fcur, fip_grid c, int ccur, int F) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y;
  if (blockIdx.x == 0 && threadIdx.x == 0 && J == 0) c.ctl->duration = f.ctl->duration;
  if (I >= c.W || J >= c.H) return;
  double sh = 0.0, sx = 0.0, sy = 0.0;
  for (int r = 0; r < F; r++)
    for (int q = 0; q < F; q++) sh += (double)f.h[fcur][(long)(F * J + r) * f.P + F * I + q];
  for (int r = 0; r < F; r++) sx += (double)f.qx[fcur][(long)(F * J + r) * f.P + F * I + F - 1];
  for (int q = 0; q < F; q++) sy += (double)f.qy[fcur][(long)(F * J + F - 1) * f.P + F * I + q];
  const double inv = 1.0 / F;
  const long o = (long)J * c.P + I;
  const float h = (float)(sh * inv * inv), qx = (float)(sx * inv), qy = (float)(sy * inv);
  c.h[ccur][o] = h; c.h[ccur ^ 1][o] = h;
  c.qx[ccur][o] = qx; c.qx[ccur ^ 1][o] = qx;
  c.qy[ccur][o] = qy; c.qy[ccur ^ 1][o] = qy;
}

// ---------------------------------------------------------------------------------------
// Input features of one coarse state into NHWC bf16 with `cstride` channels (8 for stored
// samples, 32 for a network input): z - zref, qx, qy, bed - zref, subgrid bed std, kSt,
// each scaled to order one. skip_if_halted: do nothing when the step was skipped.
__device__ __forceinline__ void feat_pixel(const ClFeat& s, int x, int y, float v[CL_FEAT]) {
  const long o = (long)y * s.P + x;
  const float b = s.bed[o];
  v[0] = (s.h[o] + b - s.zref) * 0.25f;
  v[1] = s.qx[o];
  v[2] = s.qy[o];
  v[3] = (b - s.zref) * 0.25f;
  v[4] = s.bstd[(long)y * s.W + x] * 2.0f;
  v[5] = (s.kst ? s.kst[o] : s.kst_u) * (1.0f / 30.0f);
}

extern "C" __global__ void cl_features(ClFeat s, u16* out, int cstride, const fip_ctl* ctl, int par) {
  if (ctl && ctl->halt_req[par ^ 1]) return;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= s.W || y >= s.H) return;
  float v[CL_FEAT];
  feat_pixel(s, x, y, v);
  u16* p = out + ((long)y * s.W + x) * cstride;
  for (int k = 0; k < cstride; k++) p[k] = k < CL_FEAT ? f2bf(v[k]) : (u16)0;
}

// Patches of stored samples (NHWC, 8 channels) into a batch (NHWC, 32 channels): patch b
// is ps wide and gridDim.y high, origin[3 b .. 3 b + 2] = (sample, x0, y0).
extern "C" __global__ void cl_gather(const u16* samples, long sample_elems, int W, const int* origin,
                                     int ps, u16* out) {
  const int b = blockIdx.z, y = blockIdx.y, x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= ps) return;
  const int s = origin[3 * b], x0 = origin[3 * b + 1], y0 = origin[3 * b + 2];
  const u16* src = samples + s * sample_elems + ((long)(y0 + y) * W + (x0 + x)) * 8;
  const uint4 v = *(const uint4*)src;
  uint4* dst = (uint4*)(out + (((long)b * gridDim.y + y) * ps + x) * CL_C);
  dst[0] = v;
  dst[1] = make_uint4(0, 0, 0, 0);
  dst[2] = make_uint4(0, 0, 0, 0);
  dst[3] = make_uint4(0, 0, 0, 0);
}

// ---------------------------------------------------------------------------------------
// 3x3 convolution (padding 1) of N images H x W, NHWC bf16 with 32 channels, as an
// implicit GEMM. One wavefront computes 64 pixels of one row x 32 output channels; a
// work-group holds 4 wavefronts on 4 consecutive rows. Epilogues:
//   EPI_ACT : Z = acc + bias (bf16), A = act(Z) (bf16)
//   EPI_OUT  : out = acc + bias, channels 0..1 as fp32 [pixel][2]
//   EPI_DACT: out = acc * act'(Zp) (bf16)  (backward through the previous layer)
// wp: packed weights, [mt 0..1][tap 0..8][lane 0..63][8] bf16 (A fragments).

template <int EPI, int ACT>
__global__ __launch_bounds__(256) void cl_conv(const u16* __restrict__ X, const u16* __restrict__ wp,
                                               const float* __restrict__ bias, const u16* __restrict__ Zp,
                                               u16* __restrict__ Z, u16* __restrict__ A,
                                               float* __restrict__ out, int N, int H, int W,
                                               const fip_ctl* ctl, int par) {
  if (ctl && ctl->halt_req[par ^ 1]) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int n = blockIdx.z, y = blockIdx.y * 4 + wave, x0 = blockIdx.x * 64;
  if (y >= H || n >= N) return;
  const int col = lane & 15, grp = lane >> 4;
  f32x4 acc[2][4];
  for (int m = 0; m < 2; m++)
    for (int t = 0; t < 4; t++) acc[m][t] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const u16* img = X + (long)n * H * W * CL_C;
#pragma unroll 1
  for (int tap = 0; tap < CL_TAPS; tap++) {
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int yy = y + dy;
    const bool rowok = yy >= 0 && yy < H;
    bf16x8 b[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int xx = x0 + t * 16 + col + dx;
      b[t] = (rowok && xx >= 0 && xx < W) ? ld8(img + ((long)yy * W + xx) * CL_C + 8 * grp) : zero8();
    }
    const bf16x8 a0 = ld8(wp + ((0 * CL_TAPS + tap) * 64 + lane) * 8);
    const bf16x8 a1 = ld8(wp + ((1 * CL_TAPS + tap) * 64 + lane) * 8);
#pragma unroll
    for (int t = 0; t < 4; t++) {
      acc[0][t] = mfma(a0, b[t], acc[0][t]);
      acc[1][t] = mfma(a1, b[t], acc[1][t]);
    }
  }
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int x = x0 + t * 16 + col;
      if (x >= W) continue;
      const long pix = ((long)n * H + y) * W + x;
      const int co0 = m * 16 + 4 * grp;
      if constexpr (EPI == EPI_ACT) {
        __attribute__((aligned(8))) u16 zb[4], ab[4];
        for (int r = 0; r < 4; r++) {
          const float z = acc[m][t][r] + bias[co0 + r];
          const u16 zq = f2bf(z);
          zb[r] = zq;
          ab[r] = f2bf(act<ACT>(bf2f(zq)));
        }
        *(uint2*)(Z + pix * CL_C + co0) = *(const uint2*)zb;
        *(uint2*)(A + pix * CL_C + co0) = *(const uint2*)ab;
      } else if constexpr (EPI == EPI_OUT) {
        if (co0 == 0) {
          out[pix * 2 + 0] = acc[m][t][0] + bias[0];
          out[pix * 2 + 1] = acc[m][t][1] + bias[1];
        }
      } else {
        const uint2 zr = *(const uint2*)(Zp + pix * CL_C + co0);
        const u16* zp = (const u16*)&zr;
        __attribute__((aligned(8))) u16 ob[4];
        for (int r = 0; r < 4; r++) ob[r] = f2bf(acc[m][t][r] * dact<ACT>(bf2f(zp[r])));
        *(uint2*)(Z + pix * CL_C + co0) = *(const uint2*)ob;
      }
    }
}

// ---------------------------------------------------------------------------------------
// Weight gradient of one conv layer: dW[tap][ci][co] = sum_p X[p + off(tap)][ci] dZ[p][co]
// and db[co] = sum_p dZ[p][co], as GEMMs over pixels on the matrix cores. A work-group
// takes 4-row x 64-pixel blocks (grid-stride), stages X (with a one-pixel halo) and dZ
// transposed to [channel][pixel] in LDS, and wavefront w accumulates taps w, w+4, w+8.
// Partial sums go to part[block][9*32*32 + 32] and are added by cl_reduce.
#define WG_ROWS 4
#define WG_PX 64
#define XS_W (WG_PX + 2)
#define XS_H (WG_ROWS + 2)
extern "C" __global__ __launch_bounds__(256) void cl_wgrad(const u16* __restrict__ X,
                                                          const u16* __restrict__ dZ,
                                                          float* __restrict__ part, int N, int H, int W) {
  __shared__ u16 xs[CL_C][XS_H][XS_W + 2];
  __shared__ __attribute__((aligned(16))) u16 zs[CL_C][WG_ROWS * WG_PX];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int col = lane & 15, grp = lane >> 4;
  const int segs = (W + WG_PX - 1) / WG_PX, bands = (H + WG_ROWS - 1) / WG_ROWS;
  const long nblk = (long)N * bands * segs;
  // accumulators: taps wave, wave+4, wave+8 (up to 3) x co tiles 2 x ci tiles 2
  f32x4 acc[3][2][2];
  for (int a = 0; a < 3; a++)
    for (int m = 0; m < 2; m++)
      for (int c = 0; c < 2; c++) acc[a][m][c] = (f32x4){0.f, 0.f, 0.f, 0.f};
  float dbias = 0.f;
  for (long blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int seg = (int)(blk % segs), band = (int)((blk / segs) % bands), n = (int)(blk / ((long)segs * bands));
    const int y0 = band * WG_ROWS, x0 = seg * WG_PX;
    __syncthreads();
    // stage X rows y0-1 .. y0+WG_ROWS, pixels x0-1 .. x0+WG_PX (zero outside the image)
    for (int e = threadIdx.x; e < XS_H * XS_W; e += blockDim.x) {
      const int r = e / XS_W, q = e % XS_W;
      const int yy = y0 - 1 + r, xx = x0 - 1 + q;
      const bool ok = yy >= 0 && yy < H && xx >= 0 && xx < W;
      const u16* src = X + (((long)n * H + yy) * W + xx) * CL_C;
      for (int c = 0; c < CL_C; c += 8) {
        uint4 v = ok ? *(const uint4*)(src + c) : make_uint4(0, 0, 0, 0);
        const u16* vv = (const u16*)&v;
        for (int j = 0; j < 8; j++) xs[c + j][r][q] = vv[j];
      }
    }
    for (int e = threadIdx.x; e < WG_ROWS * WG_PX; e += blockDim.x) {
      const int r = e / WG_PX, q = e % WG_PX;
      const int yy = y0 + r, xx = x0 + q;
      const bool ok = yy < H && xx < W;
      const u16* src = dZ + (((long)n * H + yy) * W + xx) * CL_C;
      for (int c = 0; c < CL_C; c += 8) {
        uint4 v = ok ? *(const uint4*)(src + c) : make_uint4(0, 0, 0, 0);
        const u16* vv = (const u16*)&v;
        for (int j = 0; j < 8; j++) zs[c + j][e] = vv[j];
      }
    }
    __syncthreads();
    if (threadIdx.x < CL_C) {
      float s = 0.f;
      for (int e = 0; e < WG_ROWS * WG_PX; e++) s += bf2f(zs[threadIdx.x][e]);
      dbias += s;
    }
    // K loop: 32 pixels per chunk (half a row)
    for (int kc = 0; kc < WG_ROWS * WG_PX / 32; kc++) {
      const int r = kc / (WG_PX / 32), px0 = (kc % (WG_PX / 32)) * 32;
      bf16x8 za[2];
      for (int m = 0; m < 2; m++) za[m] = ld8(&zs[m * 16 + col][r * WG_PX + px0 + 8 * grp]);
      for (int a = 0; a < 3; a++) {
        const int tap = wave + 4 * a;
        if (tap >= CL_TAPS) break;
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        for (int c = 0; c < 2; c++) {
          __attribute__((aligned(16))) u16 xb[8];
          const u16* src = &xs[c * 16 + col][r + 1 + dy][px0 + 8 * grp + 1 + dx];
          for (int j = 0; j < 8; j++) xb[j] = src[j];
          const bf16x8 xv = ld8(xb);
          for (int m = 0; m < 2; m++) acc[a][m][c] = mfma(za[m], xv, acc[a][m][c]);
        }
      }
    }
  }
  // D[co][ci] of tile (m, c): lane holds co = m*16 + 4 grp + rr, ci = c*16 + col
  float* pb = part + (long)blockIdx.x * (CL_TAPS * CL_C * CL_C + CL_C);
  for (int a = 0; a < 3; a++) {
    const int tap = wave + 4 * a;
    if (tap >= CL_TAPS) break;
    for (int m = 0; m < 2; m++)
      for (int c = 0; c < 2; c++)
        for (int rr = 0; rr < 4; rr++) {
          const int co = m * 16 + 4 * grp + rr, ci = c * 16 + col;
          pb[(tap * CL_C + ci) * CL_C + co] = acc[a][m][c][rr];
        }
  }
  if (threadIdx.x < CL_C) pb[CL_TAPS * CL_C * CL_C + threadIdx.x] = dbias;
}

// out[i] = sum_b part[b][i] in a fixed order (deterministic): a work-group of 16 wavefronts
// takes 64 outputs; wavefront w adds parts w, w + 16, w + 32, ... in turn, with its loads
// issued 8 at a time, and the 16 wavefront sums are added in wavefront order.
#define RED_WAVES 16
extern "C" __global__ __launch_bounds__(64 * RED_WAVES) void cl_reduce(const float* __restrict__ part, int nparts,
                                                                      int n, float* __restrict__ out) {
  __shared__ float red[RED_WAVES][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  float s = 0.f;
  if (i < n) {
    int b = wave;
    for (; b + 7 * RED_WAVES < nparts; b += 8 * RED_WAVES) {
      float v[8];
      for (int k = 0; k < 8; k++) v[k] = part[(long)(b + k * RED_WAVES) * n + i];
      for (int k = 0; k < 8; k++) s += v[k];
    }
    for (; b < nparts; b += RED_WAVES) s += part[(long)b * n + i];
  }
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && i < n) {
    float t = red[0][lane];
    for (int w = 1; w < RED_WAVES; w++) t += red[w][lane];
    out[i] = t;
  }
}

// ---------------------------------------------------------------------------------------
// Loss of a batch and its gradient. Per pixel and output channel k (qx, qy):
//   c = tanh(o) * bound * |qp|,  r = qp + c - qt,  loss = sum m r^2 / sum m
// with m = 1 on trainable faces (mask != 0). Writes dZ4 (bf16, channels 0..1, others 0)
// = d loss / d o scaled by gscale, and per-block partial sums of (m r^2, m, m (qp - qt)^2).
extern "C" __global__ void cl_loss(const float* o, const float* qp, const float* qt, const uint8_t* mask,
                                   long npix, float bound, float inv_count, u16* dz, double* part) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double s_r = 0.0, s_m = 0.0, s_p = 0.0;
  if (p < npix) {
    __attribute__((aligned(16))) u16 g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 2; k++) {
      const float m = mask[p * 2 + k] ? 1.f : 0.f;
      const float q = qp[p * 2 + k], tq = qt[p * 2 + k];
      const float th = tanhf(o[p * 2 + k]);
      const float sc = bound * fabsf(q);
      const float r = q + th * sc - tq;
      s_r += (double)(m * r * r);
      s_m += (double)m;
      s_p += (double)(m * (q - tq) * (q - tq));
      g[k] = f2bf(m * 2.f * r * inv_count * sc * (1.f - th * th));
    }
    uint4* d = (uint4*)(dz + p * CL_C);
    d[0] = *(const uint4*)g;
    d[1] = make_uint4(0, 0, 0, 0);
    d[2] = make_uint4(0, 0, 0, 0);
    d[3] = make_uint4(0, 0, 0, 0);
  }
  __shared__ double sh[3][256];
  sh[0][threadIdx.x] = s_r;
  sh[1][threadIdx.x] = s_m;
  sh[2][threadIdx.x] = s_p;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 3; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; k++) part[(long)blockIdx.x * 3 + k] = sh[k][0];
}

// The batch loss of a training step from cl_loss's partial sums (nb blocks of (m r^2, m,
// m (qp - qt)^2)): out = sum m r^2 / sum m, or 0 without a trainable face. One work-group
// of 256 threads in a fixed order, so the value does not depend on timing.
extern "C" __global__ void cl_loss_sum(const double* part, int nb, float* out) {
  __shared__ double sh[2][256];
  double r = 0.0, m = 0.0;
  for (int b = (int)threadIdx.x; b < nb; b += (int)blockDim.x) {
    r += part[(long)b * 3];
    m += part[(long)b * 3 + 1];
  }
  sh[0][threadIdx.x] = r;
  sh[1][threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[1][0] > 0.0 ? (float)(sh[0][0] / sh[1][0]) : 0.f;
}

// ---------------------------------------------------------------------------------------
// Adam on the fp32 master parameters (all layers, padded layout [layer][9*32*32 + 32]),
// keeping padded entries at 0, then the packed bf16 weights of the forward pass and of the
// backward-data pass (taps mirrored, ci and co swapped).
//   m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2; w -= lr (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps)
__device__ __forceinline__ bool cl_param_live(int layer, int idx, int nlayers) {
  if (idx >= CL_TAPS * CL_C * CL_C) {  // bias
    const int co = idx - CL_TAPS * CL_C * CL_C;
    return layer < nlayers - 1 || co < CL_OUT;
  }
  const int ci = (idx / CL_C) % CL_C, co = idx % CL_C;
  if (layer == 0 && ci >= CL_FEAT) return false;
  if (layer == nlayers - 1 && co >= CL_OUT) return false;
  return true;
}

// packed index of W[tap][ci][co] in the A-fragment layout [mt][tap][lane][8]
__device__ __forceinline__ long cl_pack_index(int tap, int ci, int co) {
  const int mt = co / 16, lane = (co % 16) + 16 * (ci / 8), j = ci % 8;
  return (((long)mt * CL_TAPS + tap) * 64 + lane) * 8 + j;
}

extern "C" __global__ void cl_pack(const float* w, int nlayers, u16* wf, u16* wb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlayers * CL_LAYER_PARAMS) return;
  const int layer = i / CL_LAYER_PARAMS, idx = i % CL_LAYER_PARAMS;
  if (idx >= CL_TAPS * CL_C * CL_C) return;
  const int tap = idx / (CL_C * CL_C), ci = (idx / CL_C) % CL_C, co = idx % CL_C;
  const u16 q = f2bf(w[i]);
  const long lw = (long)layer * 2 * CL_TAPS * 64 * 8;
  wf[lw + cl_pack_index(tap, ci, co)] = q;
  wb[lw + cl_pack_index(CL_TAPS - 1 - tap, co, ci)] = q;
}

extern "C" __global__ void cl_adam(float* w, float* m, float* v, const float* g, int nlayers, float lr,
                                   float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlayers * CL_LAYER_PARAMS) return;
  const int layer = i / CL_LAYER_PARAMS, idx = i % CL_LAYER_PARAMS;
  if (!cl_param_live(layer, idx, nlayers)) {
    w[i] = 0.f;
    return;
  }
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// One cell's state as the correction reads it: depth, east and north face discharges and
// the two network outputs.
struct ClCell {
  float h, qx, qy, ox, oy;
};

// The correction of cell (x, y)'s east and north faces (c: the cell's values, q the
// discharges to write) by the network outputs: on interior faces q += tanh(o) * bound * |q|
// unless |o| > sat (counted, the face keeps the plain step) or the domain fell back (off).
// Returns the saturated faces. It also returns the cell's mass-balance terms of the
// uncorrected state (see mass_check): v, the depth of a cell of the box [2, W-3] x
// [2, H-3], and f, the discharges the step computed through the box's faces (+ east and
// north, - west and south), each read by the thread that corrects that face, before its
// correction.
__device__ __forceinline__ unsigned correct_cell(const fip_grid& g, const ClCell& c, float* qx, float* qy, int x,
                                                 int y, float bound, float sat, bool off, double& v, double& f) {
  unsigned nsat = 0;
  v = 0.0;
  f = 0.0;
  if (x >= 2 && x <= g.W - 3 && y >= 2 && y <= g.H - 3) v = (double)c.h;
  if (x >= 1 && x <= g.W - 3 && y >= 2 && y <= g.H - 3) {
    const float q = c.qx;
    if (x == g.W - 3) f += (double)q;
    if (x == 1) f -= (double)q;
    if (fabsf(c.ox) > sat) nsat++;
    else if (!off) *qx = q + tanhf(c.ox) * bound * fabsf(q);
  }
  if (x >= 2 && x <= g.W - 3 && y >= 1 && y <= g.H - 3) {
    const float q = c.qy;
    if (y == g.H - 3) f += (double)q;
    if (y == 1) f -= (double)q;
    if (fabsf(c.oy) > sat) nsat++;
    else if (!off) *qy = q + tanhf(c.oy) * bound * fabsf(q);
  }
  return nsat;
}

// The saturated faces and the mass-balance terms of a work-group of 256 threads, written
// to its slots of `part` and `mpart` (2 doubles) in a fixed order (every thread must call
// it). The mass check adds the slots of a domain: one atomic per work-group on the
// domain's single counter would serialise, and atomic sums of doubles would depend on
// timing.
__device__ __forceinline__ void block_count(unsigned* part, double* mpart, unsigned n, double v, double f) {
  __shared__ unsigned wsum[4];
  __shared__ double wv[4], wf[4];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    n += __shfl_xor(n, m);
    v += __shfl_xor(v, m);
    f += __shfl_xor(f, m);
  }
  if ((threadIdx.x & 63) == 0) {
    wsum[threadIdx.x >> 6] = n;
    wv[threadIdx.x >> 6] = v;
    wf[threadIdx.x >> 6] = f;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    *part = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    mpart[0] = ((wv[0] + wv[1]) + wv[2]) + wv[3];
    mpart[1] = ((wf[0] + wf[1]) + wf[2]) + wf[3];
  }
}

// sum of nparts counts and mass-balance slots by one wavefront, in a fixed order (every
// lane holds the result)
__device__ __forceinline__ unsigned wave_sum_parts(const unsigned* part, const double* mpart, int nparts,
                                                   double& v, double& f) {
  unsigned t = 0;
  v = 0.0;
  f = 0.0;
  for (int i = threadIdx.x & 63; i < nparts; i += 64) {
    t += part[i];
    v += mpart[2 * i];
    f += mpart[2 * i + 1];
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    t += __shfl_xor(t, m);
    v += __shfl_xor(v, m);
    f += __shfl_xor(f, m);
  }
  return t;
}

// ---------------------------------------------------------------------------------------
// Correction of a coarse domain's current face discharges by the network output o (fp32
// [pixel][2]): on interior faces (qx: columns 1..W-3, rows 2..H-3; qy: columns 2..W-3,
// rows 1..H-3, the faces the stencil updates; the held boundary faces are never touched)
// q += tanh(o) * bound * |q|, unless |o| exceeds `sat` (the network saturates: that face
// keeps the plain step) or the domain's fallback flag is set. A corrected face moves
// water between two cells of the domain, so the volume is conserved by construction.
// One work-group per 256 columns of CL_CORRECT_ROWS rows; its count of saturated faces and
// its mass terms go to slot blockIdx.y * gridDim.x + blockIdx.x of part and mpart.
__device__ __forceinline__ unsigned correct_rows(const fip_grid& g, int cur, const float* o, float bound, float sat,
                                                 bool off, double& v, double& f) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y0 = blockIdx.y * CL_CORRECT_ROWS;
  unsigned nsat = 0;
  v = 0.0;
  f = 0.0;
  if (x >= g.W) return 0;
  // every load of the thread's rows first: the rows are independent, so no load waits for
  // the stores of the rows before it
  ClCell c[CL_CORRECT_ROWS];
#pragma unroll
  for (int r = 0; r < CL_CORRECT_ROWS; r++) {
    const int y = y0 + r;
    if (y < g.H) {
      const long i = (long)y * g.P + x;
      const float2 ov = *(const float2*)(o + ((long)y * g.W + x) * 2);
      c[r] = ClCell{g.h[cur][i], g.qx[cur][i], g.qy[cur][i], ov.x, ov.y};
    }
  }
#pragma unroll
  for (int r = 0; r < CL_CORRECT_ROWS; r++) {
    const int y = y0 + r;
    if (y >= g.H) break;
    const long i = (long)y * g.P + x;
    double cv, cf;
    nsat += correct_cell(g, c[r], g.qx[cur] + i, g.qy[cur] + i, x, y, bound, sat, off, cv, cf);
    v += cv;
    f += cf;
  }
  return nsat;
}

extern "C" __global__ __launch_bounds__(256) void cl_correct(fip_grid g, int par, const float* o, float bound,
                                                            float sat, ClMass* ms, unsigned* part, double* mpart) {
  if (g.ctl->halt_req[par ^ 1]) return;
  double v, f;
  const unsigned nsat = correct_rows(g, par ^ 1, o, bound, sat, ms->fallback != 0, v, f);
  const long slot = (long)blockIdx.y * gridDim.x + blockIdx.x;
  block_count(part + slot, mpart + 2 * slot, nsat, v, f);
  if (slot == 0 && threadIdx.x == 0) {  // the step's check is due (pending_check)
    ms->t_pend = g.ctl->duration;
    ms->pending = 1;
  }
}

// Mass balance of a coarse domain without no_data cells, checked after each step before
// its correction. The continuity update telescopes over the box of cells
// [2, W-3] x [2, H-3], which no boundary condition writes, so over the step just taken
//   V(t) - V(t - dt) = -dt dx (sum_j qx(W-3, j) - qx(1, j) + sum_i qy(i, H-3) - qy(i, 1))
// with the face discharges the step computed (stored and not yet corrected). cl_correct
// adds the box's depths and the bracket per work-group (fp64, correct_cell); the check
// adds those slots in a fixed order, compares, and sets the fallback flag when the
// relative error exceeds `limit`. t: the duration after the step.
__device__ __forceinline__ void mass_check(const fip_grid& g, double t, double vsum, double fsum, ClMass& m,
                                           double limit) {
  const double dx = g.dx;
  const double vol = vsum * dx * dx;
  if (m.valid) {
    const double pred = m.v_prev - (t - m.t_prev) * dx * fsum;
    const double rel = fabs(vol - pred) / (vol > 1e-30 ? vol : 1e-30);
    m.last_rel = rel;
    if (rel > m.max_rel) m.max_rel = rel;
    if (rel > limit && vol > 0.0) m.fallback = 1;
  }
  m.v_prev = vol;
  m.t_prev = t;
  m.valid = 1;
  m.steps = m.steps + 1;
}

// The check of the last corrected step, if it has not run yet (m.pending, set by the
// correction): one wavefront adds the saturation counts and mass terms of its nparts
// slots, then the check. It runs before the next correction, which reads the fallback
// flag: in the next step's network kernel (cl_fused_r's extra work-groups), or as
// cl_mass_check(_ens) after the correction (the tile forms) and before a status read.
__device__ __forceinline__ void pending_check(const fip_grid& g, const unsigned* part, const double* mpart,
                                              int nparts, ClMass& m, double limit) {
  if (threadIdx.x >= 64 || !m.pending) return;
  double v, f;
  const unsigned t = wave_sum_parts(part, mpart, nparts, v, f);
  if (threadIdx.x != 0) return;
  m.nsat += t;
  mass_check(g, m.t_pend, v, f, m, limit);
  m.pending = 0;
}

extern "C" __global__ void cl_mass_check(fip_grid g, const double* mpart, ClMass* ms, double limit,
                                         const unsigned* part, int nparts) {
  pending_check(g, part, mpart, nparts, *ms, limit);
}

// Ensemble forms of cl_correct / cl_mass_check: blockIdx.z (cl_mass_check_ens:
// blockIdx.x) = scenario; o holds one W x H image per scenario, part nparts slots per
// scenario.
extern "C" __global__ __launch_bounds__(256) void cl_correct_ens(const fip_ens_scn* scn, const ClDom* dom, int par,
                                                                const float* o, float bound, float sat, ClMass* ms,
                                                                unsigned* part, double* mpart) {
  const int s = blockIdx.z;
  const fip_grid& g = scn[s].g;
  if (g.ctl->halt_req[par ^ 1]) return;
  double v, f;  // work-groups past a smaller scenario's size write empty slots
  const unsigned nsat =
      correct_rows(g, par ^ 1, o + dom[s].o_off * 2, bound, sat, ms[s].fallback != 0, v, f);
  const long nparts = (long)gridDim.x * gridDim.y;
  const long slot = s * nparts + (long)blockIdx.y * gridDim.x + blockIdx.x;
  block_count(part + slot, mpart + 2 * slot, nsat, v, f);
  if (slot == s * nparts && threadIdx.x == 0) {  // the step's check is due (pending_check)
    ms[s].t_pend = g.ctl->duration;
    ms[s].pending = 1;
  }
}

extern "C" __global__ void cl_mass_check_ens(const fip_ens_scn* scn, const double* mpart, ClMass* ms, double limit,
                                             const unsigned* part, int nparts) {
  const int s = blockIdx.x;
  pending_check(scn[s].g, part + (long)s * nparts, mpart + 2L * s * nparts, nparts, ms[s], limit);
}

// ---------------------------------------------------------------------------------------
// Fused inference step of the closure for 16 x 16 coarse cells per work-group: features,
// the four convolutions (activations kept in LDS, the halo recomputed: layer l is
// evaluated on (16 + 2 (4 - l))^2 cells), the network outputs of the cells and the
// mass-balance sums, in one launch. blockIdx.z = domain. Activations outside the domain
// are 0, as the zero padding of the layer-by-layer form. The correction itself is a
// separate launch (cl_correct / cl_correct_ens): a work-group reads the discharges of a
// 4-cell halo around its tile, which other work-groups own, so correcting in place here
// would let a work-group read a neighbour's corrected faces or not, depending on timing.
// A fallback raised by the mass check after a step takes effect from the next step.
#define FT 16                 // output tile
#define FR0 (FT + 8)          // feature region
#define FR1 (FT + 6)
#define FR2 (FT + 4)
#define FR3 (FT + 2)
#define FPS 40                // LDS pixel stride of 32-channel activations (u16): conflict-free
// 8-channel packing of layer 0: K chunk c holds taps 4c .. 4c+3 (lane group = tap - 4c),
// 8 channels each; taps 9..11 are zero.
__device__ __forceinline__ long cl_pack_l1_index(int tap, int ci, int co) {
  const int mt = co / 16, chunk = tap / 4, lane = (co % 16) + 16 * (tap % 4), j = ci;
  return (((long)mt * 3 + chunk) * 64 + lane) * 8 + j;
}

extern "C" __global__ void cl_pack_l1(const float* w, u16* wl1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= CL_TAPS * 8 * CL_C) return;
  const int tap = i / (8 * CL_C), ci = (i / CL_C) % 8, co = i % CL_C;
  wl1[cl_pack_l1_index(tap, ci, co)] = f2bf(w[((long)tap * CL_C + ci) * CL_C + co]);
}

#define FNW 8   // wavefronts per work-group of cl_fused

// One layer of the fused step: wavefront w owns pixel tiles w, w + FNW, ... of the output
// region (16 pixels each, row-major over the region); for each K chunk it loads the A
// fragments once and multiplies them into all of its tiles.
template <int RIN, int ROUT, int MT, bool L1, int ACT>
__device__ __forceinline__ void fused_layer(const u16* in, u16* out, const u16* wp, const float* bias,
                                            int ox, int oy, int W, int H, float* o2) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4;
  constexpr int NPIX = ROUT * ROUT, NT = (NPIX + 15) / 16, TPW = (NT + FNW - 1) / FNW;
  constexpr int off = (RIN - ROUT) / 2;   // 1
  constexpr int NC = L1 ? 3 : CL_TAPS;    // K chunks of 32
  const int gx0 = ox - (ROUT - FT) / 2, gy0 = oy - (ROUT - FT) / 2;
  f32x4 acc[TPW][MT];
  int base[TPW];
  f32x4 bias4[MT];  // the accumulators start at the bias of their output channels
#pragma unroll
  for (int m = 0; m < MT; m++) {
    const int co0 = m * 16 + 4 * grp;
    bias4[m] = (f32x4){bias[co0], bias[co0 + 1], bias[co0 + 2], bias[co0 + 3]};
  }
#pragma unroll
  for (int i = 0; i < TPW; i++) {
#pragma unroll
    for (int m = 0; m < MT; m++) acc[i][m] = bias4[m];
    const int p = (wave + FNW * i) * 16 + col;
    const int pp = p < NPIX ? p : NPIX - 1;
    base[i] = (pp / ROUT + off) * RIN + pp % ROUT + off;   // input pixel of tap (0, 0)
  }
#pragma unroll
  for (int c = 0; c < NC; c++) {
    bf16x8 a[MT];
#pragma unroll
    for (int m = 0; m < MT; m++) a[m] = ld8(wp + ((m * NC + c) * 64 + lane) * 8);
#pragma unroll
    for (int i = 0; i < TPW; i++) {
      if (wave + FNW * i >= NT) break;
      bf16x8 b;
      if constexpr (L1) {
        const int tap = 4 * c + grp;
        const int sh = tap < CL_TAPS ? (tap / 3 - 1) * RIN + (tap % 3 - 1) : 0;
        b = tap < CL_TAPS ? ld8(in + (base[i] + sh) * 8) : zero8();
      } else {
        const int sh = (c / 3 - 1) * RIN + (c % 3 - 1);
        b = ld8(in + (base[i] + sh) * FPS + 8 * grp);
      }
#pragma unroll
      for (int m = 0; m < MT; m++) acc[i][m] = mfma(a[m], b, acc[i][m]);
    }
  }
#pragma unroll
  for (int i = 0; i < TPW; i++) {
    const int p = (wave + FNW * i) * 16 + col;
    if (p >= NPIX) continue;
    const int yy = p / ROUT, xx = p % ROUT;
    const bool inside = gx0 + xx >= 0 && gx0 + xx < W && gy0 + yy >= 0 && gy0 + yy < H;
    if (o2) {  // last layer: outputs 0 and 1, fp32, held by lane group 0
      if (grp == 0) {
        o2[p * 2] = acc[i][0][0];
        o2[p * 2 + 1] = acc[i][0][1];
      }
      continue;
    }
#pragma unroll
    for (int m = 0; m < MT; m++) {
      const int co0 = m * 16 + 4 * grp;
      // activation of the fp32 pre-activation (the layer-by-layer form rounds it to bf16 first,
      // which it stores for the backward pass; inference has no use for that rounding),
      // packed to bf16 pairs; cells outside the domain are set to 0 by a select
      const uint32_t keep = 0u - (uint32_t)inside;
      const uint32_t lo = pack_bf2(act<ACT>(acc[i][m][0]), act<ACT>(acc[i][m][1])) & keep;
      const uint32_t hi = pack_bf2(act<ACT>(acc[i][m][2]), act<ACT>(acc[i][m][3])) & keep;
      *(uint2*)(out + (yy * ROUT + xx) * FPS + co0) = make_uint2(lo, hi);
    }
  }
}

template <int ACT>
__global__ __launch_bounds__(512, 2) void cl_fused(const ClDom* dom, int par, const u16* wf,
                                                          const u16* wl1, const float* w, float* o_out) {
  __shared__ __attribute__((aligned(16))) u16 f0[FR0 * FR0 * 8];
  __shared__ __attribute__((aligned(16))) u16 b1[FR1 * FR1 * FPS];
  __shared__ __attribute__((aligned(16))) u16 b2[FR2 * FR2 * FPS];
  const int d = blockIdx.z;
  const fip_grid& g = dom[d].g;
  if (g.ctl->halt_req[par ^ 1]) return;
  const int cur = par ^ 1;
  const int W = g.W, H = g.H;
  const int ox = blockIdx.x * FT, oy = blockIdx.y * FT;
  const float* bstd = dom[d].bstd;
  const float zref = dom[d].zref;
  // features of the 24 x 24 region (the formulas of feat_pixel)
  for (int e = threadIdx.x; e < FR0 * FR0; e += blockDim.x) {
    const int gx = ox - 4 + e % FR0, gy = oy - 4 + e / FR0;
    uint4 pk = make_uint4(0, 0, 0, 0);
    if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
      const long o = (long)gy * g.P + gx;
      const float b = g.bed[o];
      const float f0v = (g.h[cur][o] + b - zref) * 0.25f;
      const float f1v = g.qx[cur][o], f2v = g.qy[cur][o];
      const float f3v = (b - zref) * 0.25f;
      const float f4v = bstd[(long)gy * W + gx] * 2.0f;
      const float f5v = (g.kst ? g.kst[o] : g.kst_uniform) * (1.0f / 30.0f);
      pk.x = (uint32_t)f2bf(f0v) | ((uint32_t)f2bf(f1v) << 16);
      pk.y = (uint32_t)f2bf(f2v) | ((uint32_t)f2bf(f3v) << 16);
      pk.z = (uint32_t)f2bf(f4v) | ((uint32_t)f2bf(f5v) << 16);
    }
    *(uint4*)(f0 + e * 8) = pk;
  }
  __syncthreads();
  const float* b0 = w + 0 * CL_LAYER_PARAMS + CL_TAPS * CL_C * CL_C;
  fused_layer<FR0, FR1, 2, true, ACT>(f0, b1, wl1, b0, ox, oy, W, H, nullptr);
  __syncthreads();
  fused_layer<FR1, FR2, 2, false, ACT>(b1, b2, wf + 1 * CL_PACKED, w + 1 * CL_LAYER_PARAMS + CL_TAPS * CL_C * CL_C, ox,
                                  oy, W, H, nullptr);
  __syncthreads();
  fused_layer<FR2, FR3, 2, false, ACT>(b2, b1, wf + 2 * CL_PACKED, w + 2 * CL_LAYER_PARAMS + CL_TAPS * CL_C * CL_C, ox,
                                  oy, W, H, nullptr);
  __syncthreads();
  float* o2 = (float*)b2;  // 16 x 16 x 2 outputs reuse the layer-2 buffer
  fused_layer<FR3, FT, 1, false, ACT>(b1, nullptr, wf + 3 * CL_PACKED, w + 3 * CL_LAYER_PARAMS + CL_TAPS * CL_C * CL_C,
                                 ox, oy, W, H, o2);
  __syncthreads();
  // the outputs of the tile's cells into o (one [H][W][2] image per domain)
  float* od = o_out + dom[d].o_off * 2;
  for (int e = threadIdx.x; e < FT * FT; e += blockDim.x) {
    const int x = ox + e % FT, y = oy + e / FT;
    if (x >= W || y >= H) continue;
    od[((long)y * W + x) * 2] = o2[e * 2];
    od[((long)y * W + x) * 2 + 1] = o2[e * 2 + 1];
  }
}

// ---------------------------------------------------------------------------------------
// Persistent form of the fused inference step: the same per-tile arithmetic (the same
// fragments summed in the same order, so the outputs are bit-identical to cl_fused), with
// the latencies of the tile form taken off the critical path:
//   * one work-group per CU takes a contiguous run of tiles; the packed weights and biases
//     of all four layers are staged in LDS once per launch, not re-read from global memory
//     by every wavefront of every tile;
//   * the input cells of tile t+1 (bed, h, qx, qy, subgrid std, kSt) are loaded into
//     registers while tile t runs layer 0, and written to the feature buffer after layer 1.
// Like cl_fused it writes the network outputs; cl_correct(_ens) applies them.
// Four barriers per tile. LDS: 52 KB of weights + 80 KB of activations (one work-group
// per CU, 8 wavefronts).
#define FP_PIX (FR0 * FR0)                    // 576 feature pixels per tile
#define FP_PER ((FP_PIX + 511) / 512)         // 2 per thread
#define FP_W1 CL_PACKED_L1                    // u16 offsets of the layers in the LDS weights
#define FP_W2 (FP_W1 + CL_PACKED)
#define FP_W3 (FP_W2 + CL_PACKED)
#define FP_WN (FP_W3 + CL_PACKED / 2)         // layer 3 uses output block 0 only

// One layer of the persistent fused step: the arithmetic of fused_layer (the same
// fragments summed in the same order, so the same bits). The wave holds the A fragments
// of every K chunk of the layer in registers and walks its pixel tiles one after the
// other; the epilogue (bias-initialised accumulators -> activation -> bf16 -> LDS) of tile i-1
// is issued between the matrix instructions of tile i, so the vector ALU work overlaps
// the matrix cores. The code is branch-free: the tile count K is a template argument
// (uniform per wave), and lanes past the region's last pixel store to a spare pixel slot
// (index NPIX) of the output buffer, or to a per-lane spare slot of o2.
template <int RIN, int ROUT, int MT, bool L1, int ACT, int K>
__device__ __forceinline__ void fp_layer_k(const u16* in, u16* out, const u16* wp, const float* bias, int ox,
                                           int oy, int W, int H, float* o2, int wave) {
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;
  constexpr int NPIX = ROUT * ROUT;
  constexpr int off = (RIN - ROUT) / 2;
  constexpr int NC = L1 ? 3 : CL_TAPS;
  const int gx0 = ox - (ROUT - FT) / 2, gy0 = oy - (ROUT - FT) / 2;
  f32x4 b4[MT];
#pragma unroll
  for (int m = 0; m < MT; m++) {
    const int co0 = m * 16 + 4 * grp;
    b4[m] = (f32x4){bias[co0], bias[co0 + 1], bias[co0 + 2], bias[co0 + 3]};
  }
  bf16x8 a[NC][MT];
#pragma unroll
  for (int c = 0; c < NC; c++)
#pragma unroll
    for (int m = 0; m < MT; m++) a[c][m] = ld8(wp + ((m * NC + c) * 64 + lane) * 8);
  const auto pix = [&](int i) { return (wave + FNW * i) * 16 + col; };
  const auto ldB = [&](int c, int i) -> bf16x8 {
    const int p = pix(i);
    const int pp = p < NPIX ? p : NPIX - 1;
    const int base = (pp / ROUT + off) * RIN + pp % ROUT + off;
    if constexpr (L1) {
      const int tap = 4 * c + grp;
      const int sh = tap < CL_TAPS ? (tap / 3 - 1) * RIN + (tap % 3 - 1) : 0;
      const bf16x8 v = ld8(in + (base + sh) * 8);
      return tap < CL_TAPS ? v : zero8();
    } else {
      const int sh = (c / 3 - 1) * RIN + (c % 3 - 1);
      return ld8(in + (base + sh) * FPS + 8 * grp);
    }
  };
  const auto epilogue = [&](int i, const f32x4 (&acc)[MT]) {
    const int p = pix(i);
    const bool valid = p < NPIX;
    if (o2) {  // last layer: outputs 0 and 1, fp32, held by lane group 0
      const int dst = (grp == 0 && valid) ? p * 2 : 1024 + 2 * lane;
      o2[dst] = acc[0][0];
      o2[dst + 1] = acc[0][1];
      return;
    }
    const int yy = p / ROUT, xx = p % ROUT;
    const bool inside = gx0 + xx >= 0 && gx0 + xx < W && gy0 + yy >= 0 && gy0 + yy < H;
    const uint32_t keep = 0u - (uint32_t)inside;
    const int px = valid ? p : NPIX;
#pragma unroll
    for (int m = 0; m < MT; m++) {
      const int co0 = m * 16 + 4 * grp;
      const uint32_t lo = pack_bf2(act<ACT>(acc[m][0]), act<ACT>(acc[m][1])) & keep;
      const uint32_t hi = pack_bf2(act<ACT>(acc[m][2]), act<ACT>(acc[m][3])) & keep;
      *(uint2*)(out + px * FPS + co0) = make_uint2(lo, hi);
    }
  };
  f32x4 prev[MT];
#pragma unroll
  for (int i = 0; i <= K; i++) {
    f32x4 acc[MT];
    if (i < K) {
#pragma unroll
      for (int m = 0; m < MT; m++) acc[m] = b4[m];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        const bf16x8 b = ldB(c, i);
#pragma unroll
        for (int m = 0; m < MT; m++) acc[m] = mfma(a[c][m], b, acc[m]);
      }
    }
    if (i > 0) epilogue(i - 1, prev);
    if (i < K) {
#pragma unroll
      for (int m = 0; m < MT; m++) prev[m] = acc[m];
    }
  }
}

// tiles are dealt round-robin, so a wave has TPW or TPW - 1 of them
template <int RIN, int ROUT, int MT, bool L1, int ACT>
__device__ __forceinline__ void fp_layer(const u16* in, u16* out, const u16* wp, const float* bias, int ox, int oy,
                                         int W, int H, float* o2) {
  constexpr int NT = (ROUT * ROUT + 15) / 16, TPW = (NT + FNW - 1) / FNW;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nt = (NT - wave + FNW - 1) / FNW;
  if (nt == TPW)
    fp_layer_k<RIN, ROUT, MT, L1, ACT, TPW>(in, out, wp, bias, ox, oy, W, H, o2, wave);
  else if constexpr (TPW > 1)
    fp_layer_k<RIN, ROUT, MT, L1, ACT, TPW - 1>(in, out, wp, bias, ox, oy, W, H, o2, wave);
}

struct FpTile {   // one tile of the persistent loop (wave-uniform)
  int d, ox, oy;
};

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ FpTile fp_tile(int t, int tx, int ty) {
  const int per = tx * ty, r = t % per;
  return FpTile{uni(t / per), uni((r % tx) * FT), uni((r / tx) * FT)};
}

// first tile in [t, end) whose domain did not halt this step (end if none)
__device__ __forceinline__ int fp_next(const ClDom* __restrict__ dom, int t, int end, int tx, int ty, int par) {
  const int per = tx * ty;
  t = uni(t);
  while (t < end) {
    const int d = uni(t / per);
    if (!dom[d].g.ctl->halt_req[par ^ 1]) break;
    t = uni((d + 1) * per);
  }
  return t < end ? t : end;
}

// What a tile needs of its domain, read once (scalar loads) one tile ahead of its use.
struct FpDom {
  const float *h, *qx, *qy, *bed, *kst, *bstd;
  float kst_u, zref;
  int W, H, P;
};

__device__ __forceinline__ FpDom fp_dom(const ClDom* __restrict__ dom, int d, int cur) {
  const ClDom& D = dom[d];
  return FpDom{D.g.h[cur], D.g.qx[cur], D.g.qy[cur], D.g.bed, D.g.kst, D.bstd, D.g.kst_uniform, D.zref,
               D.g.W, D.g.H, D.g.P};
}

struct FpCells {  // the raw inputs of this thread's feature pixels of one tile
  float b[FP_PER], h[FP_PER], qx[FP_PER], qy[FP_PER], bs[FP_PER], k[FP_PER];
};

__device__ __forceinline__ void fp_load(const FpDom& D, FpTile T, FpCells& c) {
#pragma unroll
  for (int j = 0; j < FP_PER; j++) {
    const int e = threadIdx.x + 512 * j;
    const int gx = T.ox - 4 + e % FR0, gy = T.oy - 4 + e / FR0;
    c.b[j] = c.h[j] = c.qx[j] = c.qy[j] = c.bs[j] = c.k[j] = 0.f;
    if (e < FP_PIX && gx >= 0 && gx < D.W && gy >= 0 && gy < D.H) {
      const long o = (long)gy * D.P + gx;
      c.b[j] = D.bed[o];
      c.h[j] = D.h[o];
      c.qx[j] = D.qx[o];
      c.qy[j] = D.qy[o];
      c.bs[j] = D.bstd[(long)gy * D.W + gx];
      c.k[j] = D.kst ? D.kst[o] : D.kst_u;
    }
  }
}

// features (the formulas of feat_pixel, as in cl_fused) into f0
__device__ __forceinline__ void fp_store(const FpDom& D, FpTile T, const FpCells& c, u16* f0) {
#pragma unroll
  for (int j = 0; j < FP_PER; j++) {
    const int e = threadIdx.x + 512 * j;
    if (e >= FP_PIX) continue;
    const int lx = e % FR0 - 4, ly = e / FR0 - 4;
    const int gx = T.ox + lx, gy = T.oy + ly;
    uint4 pk = make_uint4(0, 0, 0, 0);
    if (gx >= 0 && gx < D.W && gy >= 0 && gy < D.H) {
      const float b = c.b[j];
      const float f0v = (c.h[j] + b - D.zref) * 0.25f;
      const float f3v = (b - D.zref) * 0.25f;
      const float f4v = c.bs[j] * 2.0f;
      const float f5v = c.k[j] * (1.0f / 30.0f);
      pk.x = (uint32_t)f2bf(f0v) | ((uint32_t)f2bf(c.qx[j]) << 16);
      pk.y = (uint32_t)f2bf(c.qy[j]) | ((uint32_t)f2bf(f3v) << 16);
      pk.z = (uint32_t)f2bf(f4v) | ((uint32_t)f2bf(f5v) << 16);
    }
    *(uint4*)(f0 + e * 8) = pk;
  }
}

template <int ACT>
__global__ __launch_bounds__(512, 1) __attribute__((amdgpu_waves_per_eu(2, 2))) void cl_fused_p(
    const ClDom* __restrict__ dom, int par, const u16* wf, const u16* wl1, const float* w, float* o_out, int tx,
    int ty, int ntiles, int per_wg) {
  __shared__ __attribute__((aligned(16))) u16 sw[FP_WN];
  __shared__ float sb[CL_LAYERS * CL_C];
  __shared__ __attribute__((aligned(16))) u16 f0[FR0 * FR0 * 8];
  __shared__ __attribute__((aligned(16))) u16 b1[(FR1 * FR1 + 1) * FPS];  // + a spare pixel slot
  __shared__ __attribute__((aligned(16))) u16 b2[(FR2 * FR2 + 1) * FPS];
  const int t0 = uni(blockIdx.x * per_wg);
  const int end = uni(t0 + per_wg < ntiles ? t0 + per_wg : ntiles);
  const int cur = par ^ 1;
  int t = fp_next(dom, t0, end, tx, ty, par);
  if (t >= end) return;  // the whole work-group: nothing to do this step
  // weights and biases of the four layers into LDS
  for (int i = threadIdx.x; i < FP_WN / 8; i += blockDim.x) {
    // layer 0 from its 8-channel packing, layers 1..3 from the forward packing
    const u16* src = i < FP_W1 / 8 ? wl1 + 8 * i : wf + CL_PACKED + 8 * (i - FP_W1 / 8);
    *(uint4*)(sw + 8 * i) = *(const uint4*)src;
  }
  for (int i = threadIdx.x; i < CL_LAYERS * CL_C; i += blockDim.x)
    sb[i] = w[(i / CL_C) * CL_LAYER_PARAMS + CL_TAPS * CL_C * CL_C + i % CL_C];
  FpTile T = fp_tile(t, tx, ty);
  FpDom DT = fp_dom(dom, T.d, cur);
  FpCells pf;
  fp_load(DT, T, pf);
  int tn = fp_next(dom, t + 1, end, tx, ty, par);
  FpTile N = fp_tile(tn < end ? tn : t, tx, ty);
  FpDom DN = fp_dom(dom, N.d, cur);
  fp_store(DT, T, pf, f0);
  __syncthreads();
  for (;;) {
    // tile t runs the network; the inputs of tile tn are loaded meanwhile, and the
    // domain descriptor of the tile after it is read (scalar loads, used next iteration)
    if (tn < end) fp_load(DN, N, pf);  // in flight during layers 0 and 1
    const int W = DT.W, H = DT.H;
    fp_layer<FR0, FR1, 2, true, ACT>(f0, b1, sw, sb, T.ox, T.oy, W, H, nullptr);
    const int tnn = tn < end ? fp_next(dom, tn + 1, end, tx, ty, par) : end;
    const FpTile NN = fp_tile(tnn < end ? tnn : (tn < end ? tn : t), tx, ty);
    const FpDom DNN = fp_dom(dom, NN.d, cur);
    __syncthreads();
    fp_layer<FR1, FR2, 2, false, ACT>(b1, b2, sw + FP_W1, sb + CL_C, T.ox, T.oy, W, H, nullptr);
    if (tn < end) fp_store(DN, N, pf, f0);
    __syncthreads();
    fp_layer<FR2, FR3, 2, false, ACT>(b2, b1, sw + FP_W2, sb + 2 * CL_C, T.ox, T.oy, W, H, nullptr);
    __syncthreads();
    float* o2 = (float*)b2;
    fp_layer<FR3, FT, 1, false, ACT>(b1, nullptr, sw + FP_W3, sb + 3 * CL_C, T.ox, T.oy, W, H, o2);
    __syncthreads();
    const int e = threadIdx.x;
    const int x = T.ox + e % FT, y = T.oy + e / FT;
    if (e < FT * FT && x < W && y < H) {
      float* od = o_out + (dom[T.d].o_off + (long)y * W + x) * 2;
      od[0] = o2[e * 2];
      od[1] = o2[e * 2 + 1];
    }
    if (tn >= end) break;
    t = tn;
    T = N;
    DT = DN;
    tn = tnn;
    N = NN;
    DN = DNN;
  }
}

// ---------------------------------------------------------------------------------------
// Row-marching form of the fused inference step (form 2). A work-group takes a strip of
// RS = 58 output columns and a span of rows [ya, yb) of one domain and marches along its
// rows, two rows per march step. Each layer keeps its last eight rows in an LDS ring, so
// every row of every layer is computed once: no vertical halo is recomputed, and the
// horizontal halo is 3 columns per side at layer 0 (the tile forms recompute a 3-cell
// halo on all four sides). March step k loads the features of rows ya - 4 + 2k and
// ya - 3 + 2k, and layer L computes rows ya - M_L + 2k and the row after it (M_L = 8,
// 14, 18, 24), so every input row was written in an earlier half-step; two barriers per
// step. 16 wavefronts: wavefront 4 L + j runs layer L on pixel tile j (16 columns), and
// the four wavefronts on a SIMD are one of each layer. Each keeps the A fragments of its layer in registers for
// the whole march, and each B fragment it reads feeds both of its rows. Every output sums
// the same fragments in the same order as cl_fused: the same bits. cl_correct(_ens)
// applies the outputs.
#define RS 58                  // output columns per strip
#define RPX 65                 // pixel slots per ring row: 4 tiles of 16 + a spare slot
#define RFW (RS + 8)           // feature columns (4 on each side)
#define RROWS 8                // rows per ring

// Rows live in ring slot (row - ya) & 7, ya the first row of the work-group's span: at
// march step k every slot is a compile-time function of k & 3 (the march is unrolled by
// 4), so the LDS accesses below take immediate offsets from a per-lane base.
__host__ __device__ constexpr int rm_slot(int r) { return ((r % RROWS) + RROWS) % RROWS; }

// matrix part of rows y (slot S0) and y + 1 of a layer L >= 1: 9 taps, MT output blocks;
// inl = in + ((c + 1) * FPS + 8 grp) for the lane's column c. Input row y + d (d = -1..2)
// at column shift dx is one B fragment: tap (d, dx) of row y and tap (d - 1, dx) of row
// y + 1 use it, and each accumulator still takes its taps in order 0..8.
template <int MT, int S0>
__device__ __forceinline__ void rm_mfma2(const u16* inl, const bf16x8 (&a)[CL_TAPS][MT], f32x4 (&acc0)[MT],
                                         f32x4 (&acc1)[MT]) {
#pragma unroll
  for (int d = -1; d <= 2; d++) {
#pragma unroll
    for (int dx = -1; dx <= 1; dx++) {
      const bf16x8 b = ld8(inl + (rm_slot(S0 + d) * RPX + dx) * FPS);
      if (d <= 1) {
#pragma unroll
        for (int m = 0; m < MT; m++) acc0[m] = mfma(a[(d + 1) * 3 + dx + 1][m], b, acc0[m]);
      }
      if (d >= 0) {
#pragma unroll
        for (int m = 0; m < MT; m++) acc1[m] = mfma(a[d * 3 + dx + 1][m], b, acc1[m]);
      }
    }
  }
}

// layer 0's matrix part of rows y (slot S0) and y + 1: K chunk cc holds tap 4 cc + grp
// (8 feature slots each; taps past 8 are zero); off[cc][s]: the lane's feature-ring offset
// of its tap for a row in slot s
template <int S0>
__device__ __forceinline__ void rm_mfma2_l0(const u16* in, const int (&off)[3][RROWS], bool tap8,
                                            const bf16x8 (&a)[3][2], f32x4 (&acc0)[2], f32x4 (&acc1)[2]) {
#pragma unroll
  for (int cc = 0; cc < 3; cc++) {
    const bf16x8 v0 = ld8(in + off[cc][S0]);
    const bf16x8 v1 = ld8(in + off[cc][rm_slot(S0 + 1)]);
    const bool ok = cc < 2 || tap8;
    const bf16x8 b0 = ok ? v0 : zero8(), b1 = ok ? v1 : zero8();
#pragma unroll
    for (int m = 0; m < 2; m++) {
      acc0[m] = mfma(a[cc][m], b0, acc0[m]);
      acc1[m] = mfma(a[cc][m], b1, acc1[m]);
    }
  }
}

// activation epilogue of one layer row into its ring (slot S0): region column c of a region of
// width RW; lanes past the region store to the spare slot
template <int RW, int S0, int ACT>
__device__ __forceinline__ void rm_store(u16* out, int y, int c, int x, int W, int H, const f32x4 (&acc)[2],
                                         int grp) {
  const bool valid = c < RW;
  const bool inside = x >= 0 && x < W && y >= 0 && y < H;
  const uint32_t keep = 0u - (uint32_t)inside;
  const int px = valid ? c : RPX - 1;
#pragma unroll
  for (int m = 0; m < 2; m++) {
    const int co0 = m * 16 + 4 * grp;
    const uint32_t lo = pack_bf2(act<ACT>(acc[m][0]), act<ACT>(acc[m][1])) & keep;
    const uint32_t hi = pack_bf2(act<ACT>(acc[m][2]), act<ACT>(acc[m][3])) & keep;
    *(uint2*)(out + (S0 * RPX + px) * FPS + co0) = make_uint2(lo, hi);
  }
}

__device__ __forceinline__ void rm_load_a(const u16* wp, int NC, int MT, int lane, bf16x8* a) {
  for (int cc = 0; cc < NC; cc++)
    for (int m = 0; m < MT; m++) a[cc * MT + m] = ld8(wp + ((m * NC + cc) * 64 + lane) * 8);
}

#define RNW 16                 // wavefronts per work-group of cl_fused_r: 4 per layer

// features of one row for one loader lane
struct RmRow {
  float b, h, qx, qy, bs, k;
};

// One wavefront of layer L (0..3) on pixel tile j over the march steps: reads its input
// ring `in`, writes its output ring `out` (layer 3: the outputs o; it also loads the
// features of the next rows and writes them to `out`, the feature ring). Layer L computes rows ya - M + 2k and the next at step k (M = 8, 14, 18, 24).
template <int L, int ACT>
__device__ __forceinline__ void rm_march(const u16* wp, const float* bp, const u16* in, u16* out, int j, int lane,
                                         int x0, int ya, int nrows, int W, int H, float* od, const ClDom& D,
                                         int cur) {
  constexpr int NC = L == 0 ? 3 : CL_TAPS, MT = L == 3 ? 1 : 2;
  constexpr int RW = RS + 6 - 2 * L;       // region width
  constexpr int M = L == 0 ? 8 : L == 1 ? 14 : L == 2 ? 18 : 24;  // pair k: rows ya - M + 2k, +1
  constexpr int HALO = 3 - L;              // rows the next layer needs: [ya - HALO, yb + HALO)
  const int col = lane & 15, grp = lane >> 4;
  const int c = 16 * j + col;
  const int cc = c < RW ? c : RW - 1;      // clamped for the loads
  bf16x8 a[NC][MT];
  rm_load_a(wp, NC, MT, lane, &a[0][0]);
  f32x4 bias[MT];
#pragma unroll
  for (int m = 0; m < MT; m++) {
    const int co0 = m * 16 + 4 * grp;
    bias[m] = (f32x4){bp[co0], bp[co0 + 1], bp[co0 + 2], bp[co0 + 3]};
  }
  // layer 3 also loads the features: thread fc of the feature region, x = x0 + fc - 4
  const int fc = (int)threadIdx.x - 3 * 4 * 64;
  const bool loader = L == 3 && fc >= 0 && fc < RFW;
  const int fx = x0 + fc - 4;
  const bool fxin = fx >= 0 && fx < W;
  RmRow nr[2][2];  // the features of pairs k and k + 1 (loads issued two pairs ahead)
  const auto load_row = [&](RmRow& r, int yy) {
    r = RmRow{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (loader && fxin && yy >= 0 && yy < H) {
      const long o = (long)yy * D.g.P + fx;
      r.b = D.g.bed[o];
      r.h = D.g.h[cur][o];
      r.qx = D.g.qx[cur][o];
      r.qy = D.g.qy[cur][o];
      r.bs = D.bstd[(long)yy * W + fx];
      r.k = D.g.kst ? D.g.kst[o] : D.g.kst_uniform;
    }
  };
  const int nfeat = nrows + 8;             // feature rows [ya - 4, yb + 4)
  if (L == 3) {
#pragma unroll
    for (int q = 0; q < 2; q++) {
      load_row(nr[q][0], ya - 4 + 2 * q);
      load_row(nr[q][1], ya - 3 + 2 * q);
    }
  }
  const int K = (nrows + 25) / 2;          // the last pair of layer 3 holds row yb - 1
  const u16* inl = in + (cc + 1) * FPS + 8 * grp;  // layers 1..3: the lane's B-fragment base
  int l0off[3][RROWS];                               // layer 0: see rm_mfma2_l0
  if constexpr (L == 0) {
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const int tap = 4 * q + grp;
      const int dy = tap < CL_TAPS ? tap / 3 - 1 : 0, dx = tap < CL_TAPS ? tap % 3 - 1 : 0;
#pragma unroll
      for (int s0 = 0; s0 < RROWS; s0++) l0off[q][s0] = (((s0 + RROWS + dy) & (RROWS - 1)) * RFW + cc + 1 + dx) * 8;
    }
  }
  const bool tap8 = grp == 0;  // chunk 2 holds tap 8 in lane group 0 only
  // pair k of this layer: rows ya - M + 2k and the next; it is computed if it meets the
  // rows the next layer needs
  const auto active = [&](int k) { return 2 * k + 1 - M >= -HALO && 2 * k - M < nrows + HALO; };
  f32x4 acc0[MT], acc1[MT];  // a pair between its matrix instructions and its epilogue
  const auto mfma_pair = [&](auto slot) {
    constexpr int S = decltype(slot)::value;
    // a wavefront issuing matrix work goes before the epilogue wavefronts of its SIMD,
    // which fill the cycles the matrix pipe leaves
    __builtin_amdgcn_s_setprio(2);
#pragma unroll
    for (int m = 0; m < MT; m++) acc0[m] = acc1[m] = bias[m];
    if constexpr (L == 0)
      rm_mfma2_l0<S>(in, l0off, tap8, a, acc0, acc1);
    else
      rm_mfma2<MT, S>(inl, a, acc0, acc1);
    __builtin_amdgcn_s_setprio(0);
  };
  const auto epi_pair = [&](auto slot, int k) {
    constexpr int S = decltype(slot)::value;
    const int y = ya - M + 2 * k;
    if constexpr (L < 3) {
      rm_store<RW, S, ACT>(out, y, c, x0 + c - HALO, W, H, acc0, grp);
      rm_store<RW, rm_slot(S + 1), ACT>(out, y + 1, c, x0 + c - HALO, W, H, acc1, grp);
    } else {  // outputs 0 and 1 (lane group 0) of rows y and y + 1 that the span owns
      const int x = x0 + c;
      if (grp == 0 && c < RS && x < W) {
        if (y >= ya && y < ya + nrows) {
          od[((long)y * W + x) * 2] = acc0[0][0];
          od[((long)y * W + x) * 2 + 1] = acc0[0][1];
        }
        if (y + 1 >= ya && y + 1 < ya + nrows) {
          od[((long)(y + 1) * W + x) * 2] = acc1[0][0];
          od[((long)(y + 1) * W + x) * 2 + 1] = acc1[0][1];
        }
      }
    }
  };
  // the loader lanes of layer 3: features of rows ya - 4 + 2k and the next into the
  // feature ring, their mass terms, and the loads of pair k + 2
  const auto feat_pair = [&](auto phase, int k) {
    constexpr int PH = decltype(phase)::value;
    const int yF = ya - 4 + 2 * k;                   // slot (yF - ya) & 7 = rm_slot(2 PH - 4)
    if (!loader) return;
#pragma unroll
    for (int r = 0; r < 2; r++) {
      const int yr = yF + r;
      if (2 * k + r < nfeat) {  // features of row yr (the formulas of feat_pixel)
        const RmRow& w = nr[PH & 1][r];
        uint4 pk = make_uint4(0, 0, 0, 0);
        if (fxin && yr >= 0 && yr < H) {
          const float f0v = (w.h + w.b - D.zref) * 0.25f;
          const float f3v = (w.b - D.zref) * 0.25f;
          const float f4v = w.bs * 2.0f;
          const float f5v = w.k * (1.0f / 30.0f);
          pk.x = (uint32_t)f2bf(f0v) | ((uint32_t)f2bf(w.qx) << 16);
          pk.y = (uint32_t)f2bf(w.qy) | ((uint32_t)f2bf(f3v) << 16);
          pk.z = (uint32_t)f2bf(f4v) | ((uint32_t)f2bf(f5v) << 16);
        }
        *(uint4*)(out + (rm_slot(2 * PH - 4 + r) * RFW + fc) * 8) = pk;  // the feature ring
      }
    }
    if (2 * k + 4 < nfeat) {  // pair k + 2: HBM latency hides behind two march steps
      load_row(nr[PH & 1][0], yF + 4);
      load_row(nr[PH & 1][1], yF + 5);
    }
  };
  // A march step is two halves, each closed by a barrier. Layers 1 and 3 issue the matrix
  // instructions of pair k in the first half, layers 0 and 2 in the second; layer 1
  // stores its pair in the second half, layers 0 and 2 theirs in the first half of the
  // next step, layer 3 at once. So on every SIMD the matrix work of two wavefronts meets
  // the epilogues of the other two. The row offsets M make every input row stored in an
  // earlier half-step, and the 8-row rings hold each row until its last reader is done.
  const auto step = [&](auto phase, int k) {
    constexpr int PH = decltype(phase)::value;       // k & 3
    using Cur = std::integral_constant<int, rm_slot(2 * PH - M)>;       // slot of pair k
    using Prev = std::integral_constant<int, rm_slot(2 * PH - 2 - M)>;  // slot of pair k - 1
    if constexpr (L == 1 || L == 3) {
      if (active(k)) {
        mfma_pair(Cur{});
        if constexpr (L == 3) epi_pair(Cur{}, k);
      }
    } else {
      if (k >= 1 && active(k - 1)) epi_pair(Prev{}, k - 1);
    }
    __syncthreads();
    if constexpr (L == 0 || L == 2) {
      if (active(k)) mfma_pair(Cur{});
    } else if constexpr (L == 1) {
      if (active(k)) epi_pair(Cur{}, k);
    } else {
      feat_pair(phase, k);
    }
    __syncthreads();
  };
  for (int k0 = 0; k0 < K; k0 += 4) {  // K is the same in every wavefront of the work-group
    step(std::integral_constant<int, 0>{}, k0);
    if (k0 + 1 < K) step(std::integral_constant<int, 1>{}, k0 + 1);
    if (k0 + 2 < K) step(std::integral_constant<int, 2>{}, k0 + 2);
    if (k0 + 3 < K) step(std::integral_constant<int, 3>{}, k0 + 3);
  }
}

// Work-groups nitems.. (one per domain, when mc is given) run the pending mass check of the
// step before (pending_check), which the correction of this step waits for.
template <int ACT>
__global__ __launch_bounds__(1024, 1) void cl_fused_r(const ClDom* __restrict__ dom, int par,
                                                             const u16* wf, const u16* wl1, const float* w,
                                                             float* o_out, const int4* __restrict__ items,
                                                             int nitems, ClMassCheck mc) {
  __shared__ __attribute__((aligned(16))) u16 fring[RROWS * RFW * 8];
  __shared__ __attribute__((aligned(16))) u16 ring1[RROWS * RPX * FPS];
  __shared__ __attribute__((aligned(16))) u16 ring2[RROWS * RPX * FPS];
  __shared__ __attribute__((aligned(16))) u16 ring3[RROWS * RPX * FPS];
  if ((int)blockIdx.x >= nitems) {
    const int d = blockIdx.x - nitems;
    pending_check(dom[d].g, mc.part + (long)d * mc.nparts, mc.mpart + 2L * d * mc.nparts, mc.nparts, mc.ms[d],
                  mc.limit);
    return;
  }
  const int4 it = items[blockIdx.x];  // (domain, first column, first row, end row)
  const int d = uni(it.x);
  const ClDom& D = dom[d];
  if (D.g.ctl->halt_req[par ^ 1]) return;  // the whole work-group
  const int cur = par ^ 1;
  const int W = D.g.W, H = D.g.H;
  const int x0 = uni(it.y), ya = uni(it.z), yb = uni(it.w);
  const int nrows = yb - ya;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int layer = wave >> 2, j = wave & 3;
  float* od = o_out + D.o_off * 2;
  const float* bias = w + CL_TAPS * CL_C * CL_C;  // + layer * CL_LAYER_PARAMS
  if (layer == 0)
    rm_march<0, ACT>(wl1, bias, fring, ring1, j, lane, x0, ya, nrows, W, H, od, D, cur);
  else if (layer == 1)
    rm_march<1, ACT>(wf + CL_PACKED, bias + CL_LAYER_PARAMS, ring1, ring2, j, lane, x0, ya, nrows, W, H, od, D, cur);
  else if (layer == 2)
    rm_march<2, ACT>(wf + 2 * CL_PACKED, bias + 2 * CL_LAYER_PARAMS, ring2, ring3, j, lane, x0, ya, nrows, W, H, od, D,
                cur);
  else
    rm_march<3, ACT>(wf + 3 * CL_PACKED, bias + 3 * CL_LAYER_PARAMS, ring3, fring, j, lane, x0, ya, nrows, W, H, od, D,
                cur);
}

// ---------------------------------------------------------------------------------------
// Targets of a training sample: the fine domain's current face discharges restricted as
// in cl_restrict, [pixel][2] fp32, and the mask of trainable faces (the ones cl_correct
// corrects).
extern "C" __global__ void cl_restrict_flux(fip_grid f, int fcur, int F, int Wc, int Hc, float* qt,
                                            uint8_t* mask) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y;
  if (I >= Wc || J >= Hc) return;
  double sx = 0.0, sy = 0.0;
  for (int r = 0; r < F; r++) sx += (double)f.qx[fcur][(long)(F * J + r) * f.P + F * I + F - 1];
  for (int q = 0; q < F; q++) sy += (double)f.qy[fcur][(long)(F * J + F - 1) * f.P + F * I + q];
  const long p = (long)J * Wc + I;
  qt[p * 2] = (float)(sx / F);
  qt[p * 2 + 1] = (float)(sy / F);
  mask[p * 2] = (I >= 1 && I <= Wc - 3 && J >= 2 && J <= Hc - 3) ? 1 : 0;
  mask[p * 2 + 1] = (I >= 2 && I <= Wc - 3 && J >= 1 && J <= Hc - 3) ? 1 : 0;
}

// The current face discharges of a domain, [pixel][2] fp32 (the plain step's result).
extern "C" __global__ void cl_flux(fip_grid g, int cur, float* q) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= g.W || y >= g.H) return;
  const long p = (long)y * g.W + x, i = (long)y * g.P + x;
  q[p * 2] = g.qx[cur][i];
  q[p * 2 + 1] = g.qy[cur][i];
}

// Patches of the per-pixel sample arrays (qp, qt fp32 [pixel][2], mask u8 [pixel][2]).
extern "C" __global__ void cl_gather_targets(const float* qp, const float* qt, const uint8_t* mask,
                                             long sample_pixels, int W, const int* origin, int ps,
                                             float* bqp, float* bqt, uint8_t* bmask) {
  const int b = blockIdx.z, y = blockIdx.y, x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= ps) return;
  const int s = origin[3 * b], x0 = origin[3 * b + 1], y0 = origin[3 * b + 2];
  const long src = s * sample_pixels + (long)(y0 + y) * W + (x0 + x);
  const long dst = ((long)b * gridDim.y + y) * ps + x;
  for (int k = 0; k < 2; k++) {
    bqp[dst * 2 + k] = qp[src * 2 + k];
    bqt[dst * 2 + k] = qt[src * 2 + k];
    bmask[dst * 2 + k] = mask[src * 2 + k];
  }
}

// fp32 features [pixel][6] (host-provided, for tests) into NHWC bf16 with 32 channels.
extern "C" __global__ void cl_feat_from_float(const float* f, long npix, u16* out) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  for (int k = 0; k < CL_C; k++) out[p * CL_C + k] = k < CL_FEAT ? f2bf(f[p * CL_FEAT + k]) : (u16)0;
}

// ---------------------------------------------------------------------------------------
// launchers (host side of this translation unit)
template <int ACT>
void launch_conv(int epi, dim3 grid, dim3 block, const u16* X, const u16* wp, const float* bias, const u16* Zp,
                 u16* Z, u16* A, float* out, int N, int H, int W, const fip_ctl* ctl, int par, hipStream_t st) {
  switch (epi) {
    case EPI_ACT:
      hipLaunchKernelGGL((cl_conv<EPI_ACT, ACT>), grid, block, 0, st, X, wp, bias, Zp, Z, A, out, N, H, W, ctl, par);
      break;
    case EPI_OUT:
      hipLaunchKernelGGL((cl_conv<EPI_OUT, ACT>), grid, block, 0, st, X, wp, bias, Zp, Z, A, out, N, H, W, ctl, par);
      break;
    default:
      hipLaunchKernelGGL((cl_conv<EPI_DACT, ACT>), grid, block, 0, st, X, wp, bias, Zp, Z, A, out, N, H, W, ctl,
                         par);
      break;
  }
}

hipError_t cl_launch_conv(int epi, int act, const u16* X, const u16* wp, const float* bias, const u16* Zp, u16* Z,
                          u16* A, float* out, int N, int H, int W, const fip_ctl* ctl, int par,
                          hipStream_t st) {
  const dim3 grid((W + 63) / 64, (H + 3) / 4, N), block(256);
  if (act == CL_ACT_HSILU)
    launch_conv<CL_ACT_HSILU>(epi, grid, block, X, wp, bias, Zp, Z, A, out, N, H, W, ctl, par, st);
  else
    launch_conv<CL_ACT_GELU>(epi, grid, block, X, wp, bias, Zp, Z, A, out, N, H, W, ctl, par, st);
  return hipGetLastError();
}

#define CL_WRAP(name, params, args, grid, block)                          \
  hipError_t name##_launch params {                                      \
    hipLaunchKernelGGL(name, grid, block, 0, st, args);                   \
    return hipGetLastError();                                             \
  }
#define P(...) __VA_ARGS__

CL_WRAP(cl_mfma_probe, (const u16* A, const u16* B, float* D, hipStream_t st), P(A, B, D), dim3(1), dim3(64))
CL_WRAP(cl_restrict, (fip_grid f, int fcur, fip_grid c, int ccur, int F, hipStream_t st),
        P(f, fcur, c, ccur, F), dim3((c.W + 255) / 256, c.H), dim3(256))
CL_WRAP(cl_restrict_flux, (fip_grid f, int fcur, int F, int Wc, int Hc, float* qt, uint8_t* mask, hipStream_t st),
        P(f, fcur, F, Wc, Hc, qt, mask), dim3((Wc + 255) / 256, Hc), dim3(256))
CL_WRAP(cl_flux, (fip_grid g, int cur, float* q, hipStream_t st), P(g, cur, q), dim3((g.W + 255) / 256, g.H),
        dim3(256))
CL_WRAP(cl_features, (ClFeat s, u16* out, int cstride, const fip_ctl* ctl, int par, hipStream_t st),
        P(s, out, cstride, ctl, par), dim3((s.W + 255) / 256, s.H), dim3(256))
CL_WRAP(cl_gather, (const u16* samples, long sample_elems, int W, const int* origin, int ps, int ph, u16* out,
                    int nb, hipStream_t st),
        P(samples, sample_elems, W, origin, ps, out), dim3((ps + 127) / 128, ph, nb), dim3(128))
CL_WRAP(cl_gather_targets, (const float* qp, const float* qt, const uint8_t* mask, long sample_pixels, int W,
                            const int* origin, int ps, int ph, float* bqp, float* bqt, uint8_t* bmask, int nb,
                            hipStream_t st),
        P(qp, qt, mask, sample_pixels, W, origin, ps, bqp, bqt, bmask), dim3((ps + 127) / 128, ph, nb),
        dim3(128))
CL_WRAP(cl_feat_from_float, (const float* f, long npix, u16* out, hipStream_t st), P(f, npix, out),
        dim3((unsigned)((npix + 255) / 256)), dim3(256))
CL_WRAP(cl_wgrad, (const u16* X, const u16* dZ, float* part, int N, int H, int W, int nblocks, hipStream_t st),
        P(X, dZ, part, N, H, W), dim3(nblocks), dim3(256))
CL_WRAP(cl_reduce, (const float* part, int nparts, int n, float* out, hipStream_t st), P(part, nparts, n, out),
        dim3((n + 63) / 64), dim3(64 * RED_WAVES))
CL_WRAP(cl_loss, (const float* o, const float* qp, const float* qt, const uint8_t* mask, long npix, float bound,
                  float inv_count, u16* dz, double* part, hipStream_t st),
        P(o, qp, qt, mask, npix, bound, inv_count, dz, part), dim3((unsigned)((npix + 255) / 256)), dim3(256))
CL_WRAP(cl_pack, (const float* w, int nlayers, u16* wf, u16* wb, hipStream_t st), P(w, nlayers, wf, wb),
        dim3((nlayers * CL_LAYER_PARAMS + 255) / 256), dim3(256))
CL_WRAP(cl_loss_sum, (const double* part, int nb, float* out, hipStream_t st), P(part, nb, out), dim3(1), dim3(256))
CL_WRAP(cl_adam, (float* w, float* m, float* v, const float* g, int nlayers, float lr, float b1, float b2,
                  float eps, float c1, float c2, hipStream_t st),
        P(w, m, v, g, nlayers, lr, b1, b2, eps, c1, c2), dim3((nlayers * CL_LAYER_PARAMS + 255) / 256), dim3(256))
CL_WRAP(cl_correct, (fip_grid g, int par, const float* o, float bound, float sat, ClMass* ms, unsigned* part,
                     double* mpart, hipStream_t st),
        P(g, par, o, bound, sat, ms, part, mpart),
        dim3((g.W + 255) / 256, (g.H + CL_CORRECT_ROWS - 1) / CL_CORRECT_ROWS), dim3(256))
CL_WRAP(cl_mass_check, (fip_grid g, const double* mpart, ClMass* ms, double limit, const unsigned* part, int nparts,
                        hipStream_t st),
        P(g, mpart, ms, limit, part, nparts), dim3(1), dim3(64))
CL_WRAP(cl_correct_ens, (const fip_ens_scn* scn, const ClDom* dom, int par, const float* o, float bound, float sat,
                         ClMass* ms, unsigned* part, double* mpart, int W, int H, int n, hipStream_t st),
        P(scn, dom, par, o, bound, sat, ms, part, mpart),
        dim3((W + 255) / 256, (H + CL_CORRECT_ROWS - 1) / CL_CORRECT_ROWS, n), dim3(256))
CL_WRAP(cl_mass_check_ens, (const fip_ens_scn* scn, int n, const double* mpart, ClMass* ms, double limit,
                            const unsigned* part, int nparts, hipStream_t st),
        P(scn, mpart, ms, limit, part, nparts), dim3(n), dim3(64))
CL_WRAP(cl_pack_l1, (const float* w, u16* wl1, hipStream_t st), P(w, wl1), dim3((CL_TAPS * 8 * CL_C + 255) / 256),
        dim3(256))

std::vector<int4> cl_fused_plan(const int* W, const int* H, int ndom, int cus) {
  // strips of RS columns, spans of rows chosen so that the work-groups (one per CU at a
  // time) finish in the fewest march steps, (span + 25) / 2 each
  int hmax = 1;
  for (int d = 0; d < ndom; d++) hmax = H[d] > hmax ? H[d] : hmax;
  int best = hmax;
  long best_cost = -1;
  for (int span = 16; span < hmax + 16; span += 16) {
    const int sp = span < hmax ? span : hmax;
    long items = 0;
    for (int d = 0; d < ndom; d++) items += (long)((W[d] + RS - 1) / RS) * ((H[d] + sp - 1) / sp);
    const long rounds = (items + cus - 1) / cus;
    const long cost = rounds * ((sp + 25) / 2);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = sp;
    }
    if (sp == hmax) break;
  }
  std::vector<int4> out;
  for (int d = 0; d < ndom; d++)
    for (int ya = 0; ya < H[d]; ya += best)
      for (int x0 = 0; x0 < W[d]; x0 += RS) out.push_back(make_int4(d, x0, ya, ya + best < H[d] ? ya + best : H[d]));
  return out;
}

template <int ACT>
hipError_t fused_launch(const ClDom* dom, int ndom, int W, int H, int par, const u16* wf, const u16* wl1,
                        const float* w, float* o, hipStream_t st, int form, const int4* items, int nitems,
                        const ClMassCheck* mc) {
  const int tx = (W + FT - 1) / FT, ty = (H + FT - 1) / FT;
  if (form == 0) {
    hipLaunchKernelGGL(cl_fused<ACT>, dim3(tx, ty, ndom), dim3(512), 0, st, dom, par, wf, wl1, w, o);
    return hipGetLastError();
  }
  if (form == 2) {
    if (!items || nitems < 1) return hipErrorInvalidValue;
    const ClMassCheck m = mc ? *mc : ClMassCheck{};
    hipLaunchKernelGGL(cl_fused_r<ACT>, dim3(nitems + (mc ? ndom : 0)), dim3(64 * RNW), 0, st, dom, par, wf, wl1,
                       w, o, items, nitems, m);
    return hipGetLastError();
  }
  static int cus[64] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  if (!cus[dev]) {
    e = hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
  }
  // persistent: one work-group per CU, contiguous runs of tiles
  const int ntiles = tx * ty * ndom;
  const int per = (ntiles + cus[dev] - 1) / cus[dev];
  const int groups = (ntiles + per - 1) / per;
  hipLaunchKernelGGL(cl_fused_p<ACT>, dim3(groups), dim3(512), 0, st, dom, par, wf, wl1, w, o, tx, ty, ntiles, per);
  return hipGetLastError();
}

hipError_t cl_fused_launch(const ClDom* dom, int ndom, int W, int H, int par, const u16* wf, const u16* wl1,
                           const float* w, float* o, hipStream_t st, int form, const int4* items, int nitems,
                           int act, const ClMassCheck* mc) {
  if (act == CL_ACT_HSILU)
    return fused_launch<CL_ACT_HSILU>(dom, ndom, W, H, par, wf, wl1, w, o, st, form, items, nitems, mc);
  return fused_launch<CL_ACT_GELU>(dom, ndom, W, H, par, wf, wl1, w, o, st, form, items, nitems, mc);
}
