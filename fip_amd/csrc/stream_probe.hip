// Memory ceiling of the dense stencil's traffic on this device, without its arithmetic.
//
// The stencil step moves 29 B per cell on the 5667 x 6500 workload: h, qx, qy and bed are
// read as floats, the kSt class as one byte, and h, qx, qy are written. This program moves
// the same bytes in three ways and prints the bandwidth of each:
//   march     the stencil's own access pattern: one wavefront per 60-column tile column
//             (64 lanes: 4 halo columns), marching a span of rows (two warm-up rows), one
//             4-byte (or 1-byte) access per lane and field per row; stores skip the halo
//             lanes and warm-up rows;
//   march/8   the same with the 8 wavefronts of a work-group on adjacent tile columns,
//             stepping through the rows in step (a barrier every second row), as the fast
//             dense kernel does;
//   march/lds the march with LDS-staged rows instead of per-lane loads: the work-group's
//             threads load each row of its 8 tile columns (484 columns with the halo)
//             once into LDS, and every lane reads its column and both neighbours from
//             LDS (two barriers per row) in place of the DPP lane shifts;
//   flat      the same byte counts as one flat stream: 16-byte accesses per lane
//             (4 bytes per lane for the class field), grid-stride;
//   flat/nt   the flat stream with non-temporal (streaming) stores;
//   aos       the same cells with h, qx, qy, bed interleaved as one float4 per cell: one
//             16-byte read and one 16-byte write per cell (bed rewritten) plus the class
//             byte, 33 B per cell, with plain or non-temporal stores;
//   copy      a plain 16-byte-per-lane copy of one field (1 read + 1 write);
//   read      the flat stream's reads alone (4 floats and the class byte per cell);
//   write     the flat stream's writes alone (3 floats per cell);
//   copy4     four fields copied at once (4 reads + 4 writes: eight streams at 1:1);
//   stagger   the flat stream with the eight fields in one allocation, field k starting
//             k * (field size + s) bytes in, for a stagger s;
//   seg/soa   one wavefront per 64-column row segment, 4 bytes per lane and field, the
//             fields in separate arrays: eight streams (h, qx, qy, bed, class read;
//             h, qx, qy written);
//   seg/aosoa the same accesses with h, qx, qy of a segment contiguous (64 h, then 64 qx,
//             then 64 qy: 768 B), bed and the class in their own arrays: four streams
//             (the h|qx|qy array and bed and class read, an h|qx|qy array written).
//             The segments cover the 5696-column pitch (0.5 % more cells than the grid).
// Every access stays inside its allocation: rows are below H, columns are clamped to
// [0, W - 1], flat indices are below W * H / 4, and the staggered fields end inside the
// allocation (checked on the host).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kW = 5667, kH = 6500, kP = 5696;  // pitch: a multiple of 64 floats
constexpr int kIn = 4, kOut = 3;

struct Fields {
  const float* in[kIn];
  const unsigned char* cls;
  float* out[kOut];
};

template <bool STEP>
__global__ __launch_bounds__(512) void march(Fields f, int span, int tiles_x, int groups) {
  const int wpb = blockDim.x >> 6, lane = threadIdx.x & 63;
  const int sr = blockIdx.x / groups;
  const int tx = (blockIdx.x % groups) * wpb + (threadIdx.x >> 6);
  int c = tx * 60 + lane - 2;
  const bool own = lane >= 2 && lane < 62 && c >= 0 && c < kW;
  c = c < 0 ? 0 : (c >= kW ? kW - 1 : c);
  const int gfirst = sr * span;
  const int r0 = gfirst >= 2 ? gfirst - 2 : 0;
  const int r1 = gfirst + span < kH ? gfirst + span : kH;
  float acc = 0.f;
  for (int r = r0; r < r1; ++r) {
    if (STEP && ((r - r0) & 1) == 0) __builtin_amdgcn_s_barrier();
    const size_t o = (size_t)r * kP + c;
    float s = f.cls[o];
#pragma unroll
    for (int k = 0; k < kIn; k++) s += f.in[k][o];
    acc = acc * 0.5f + s;
    if (own && r >= gfirst) {
#pragma unroll
      for (int k = 0; k < kOut; k++) f.out[k][o] = acc + (float)k;
    }
  }
}

// LDS-staged march: 8 wavefronts of 60-column tiles, rows staged through LDS.
__global__ __launch_bounds__(512) void march_lds(Fields f, int span, int groups) {
  constexpr int kCols = 8 * 60 + 4;
  __shared__ float sh[kIn][kCols];
  __shared__ float scl[kCols];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
  const int sr = blockIdx.x / groups;
  const int c0 = (blockIdx.x % groups) * 8 * 60 - 2;  // column of LDS slot 0
  const int my = wave * 60 + lane;                     // this lane's slot, 0 .. kCols - 1
  const int c = c0 + my;
  const bool own = lane >= 2 && lane < 62 && c >= 0 && c < kW;
  const int cst = c < 0 ? 0 : (c >= kW ? kW - 1 : c);
  int cl = c0 + t;
  cl = cl < 0 ? 0 : (cl >= kW ? kW - 1 : cl);
  const int lw = my > 0 ? my - 1 : 0, le = my < kCols - 1 ? my + 1 : kCols - 1;
  const int gfirst = sr * span;
  const int r0 = gfirst >= 2 ? gfirst - 2 : 0;
  const int r1 = gfirst + span < kH ? gfirst + span : kH;
  float acc = 0.f;
  for (int r = r0; r < r1; ++r) {
    if (t < kCols) {
      const size_t o = (size_t)r * kP + cl;
#pragma unroll
      for (int k = 0; k < kIn; k++) sh[k][t] = f.in[k][o];
      scl[t] = f.cls[o];
    }
    __syncthreads();
    float s = scl[my];
#pragma unroll
    for (int k = 0; k < kIn; k++) s += sh[k][my] + 0.5f * (sh[k][lw] - sh[k][le]);
    __syncthreads();
    acc = acc * 0.5f + s;
    if (own && r >= gfirst) {
      const size_t o = (size_t)r * kP + cst;
#pragma unroll
      for (int k = 0; k < kOut; k++) f.out[k][o] = acc + (float)k;
    }
  }
}

__global__ __launch_bounds__(256) void flat(Fields f, size_t n4) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
    const uchar4 cl = reinterpret_cast<const uchar4*>(f.cls)[i];
    float4 s = make_float4(cl.x, cl.y, cl.z, cl.w);
#pragma unroll
    for (int k = 0; k < kIn; k++) {
      const float4 v = reinterpret_cast<const float4*>(f.in[k])[i];
      s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
    }
#pragma unroll
    for (int k = 0; k < kOut; k++) reinterpret_cast<float4*>(f.out[k])[i] = s;
  }
}

template <bool NT>
__global__ __launch_bounds__(256) void flat_st(Fields f, size_t n4) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
    const uchar4 cl = reinterpret_cast<const uchar4*>(f.cls)[i];
    float4 s = make_float4(cl.x, cl.y, cl.z, cl.w);
#pragma unroll
    for (int k = 0; k < kIn; k++) {
      const float4 v = reinterpret_cast<const float4*>(f.in[k])[i];
      s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
    }
#pragma unroll
    for (int k = 0; k < kOut; k++) {
      float4* d = reinterpret_cast<float4*>(f.out[k]) + i;
      if (NT) {
        __builtin_nontemporal_store(s.x, &d->x);
        __builtin_nontemporal_store(s.y, &d->y);
        __builtin_nontemporal_store(s.z, &d->z);
        __builtin_nontemporal_store(s.w, &d->w);
      } else {
        *d = s;
      }
    }
  }
}

// n cells: a[i] = (h, qx, qy, bed), cls[i]; b[i] = (h', qx', qy', bed)
template <bool NT>
__global__ __launch_bounds__(256) void aos(const float4* a, const unsigned char* cls, float4* b, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float4 v = a[i];
    const float c = cls[i];
    const float4 r = make_float4(v.x + c, v.y * c, v.z - c, v.w);
    if (NT) {
      __builtin_nontemporal_store(r.x, &b[i].x);
      __builtin_nontemporal_store(r.y, &b[i].y);
      __builtin_nontemporal_store(r.z, &b[i].z);
      __builtin_nontemporal_store(r.w, &b[i].w);
    } else {
      b[i] = r;
    }
  }
}

// One wavefront per 64-column segment (nseg segments, grid-stride), fields in separate
// arrays of 64 * nseg elements.
template <bool NT>
__global__ __launch_bounds__(256) void seg_soa(Fields f, size_t nseg) {
  const int lane = threadIdx.x & 63;
  const size_t waves = ((size_t)gridDim.x * blockDim.x) >> 6;
  for (size_t sg = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; sg < nseg; sg += waves) {
    const size_t o = sg * 64 + lane;
    const float h = f.in[0][o], qx = f.in[1][o], qy = f.in[2][o], b = f.in[3][o];
    const float c = f.cls[o];
    const float r[3] = {h + b + c, qx * b, qy - c};
#pragma unroll
    for (int k = 0; k < kOut; k++) {
      if (NT) __builtin_nontemporal_store(r[k], &f.out[k][o]);
      else f.out[k][o] = r[k];
    }
  }
}

// The same with h, qx, qy of segment sg at hqq[192 sg + 64 k + lane] (k = 0, 1, 2), bed and
// the class at [64 sg + lane], the result at out[192 sg + 64 k + lane].
template <bool NT>
__global__ __launch_bounds__(256) void seg_aosoa(const float* hqq, const float* bed, const unsigned char* cls,
                                                 float* out, size_t nseg) {
  const int lane = threadIdx.x & 63;
  const size_t waves = ((size_t)gridDim.x * blockDim.x) >> 6;
  for (size_t sg = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; sg < nseg; sg += waves) {
    const float* src = hqq + sg * 192 + lane;
    const float h = src[0], qx = src[64], qy = src[128];
    const float b = bed[sg * 64 + lane];
    const float c = cls[sg * 64 + lane];
    const float r[3] = {h + b + c, qx * b, qy - c};
    float* d = out + sg * 192 + lane;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      if (NT) __builtin_nontemporal_store(r[k], d + 64 * k);
      else d[64 * k] = r[k];
    }
  }
}

__global__ __launch_bounds__(256) void rd(Fields f, size_t n4, float* sink) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
    const uchar4 cl = reinterpret_cast<const uchar4*>(f.cls)[i];
    s.x += cl.x;
    s.y += cl.w;
#pragma unroll
    for (int k = 0; k < kIn; k++) {
      const float4 v = reinterpret_cast<const float4*>(f.in[k])[i];
      s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
    }
  }
  if (s.x + s.y + s.z + s.w == 1234.5f) sink[0] = s.x;  // keeps the loads; never true here
}

__global__ __launch_bounds__(256) void wr(Fields f, size_t n4) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
    const float4 v = make_float4((float)i, 1.f, 2.f, 3.f);
#pragma unroll
    for (int k = 0; k < kOut; k++) reinterpret_cast<float4*>(f.out[k])[i] = v;
  }
}

// in[0..2] -> out[0..2] and in[3] -> o4
__global__ __launch_bounds__(256) void copy4(Fields f, float* o4, size_t n4) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
#pragma unroll
    for (int k = 0; k < kOut; k++)
      reinterpret_cast<float4*>(f.out[k])[i] = reinterpret_cast<const float4*>(f.in[k])[i];
    reinterpret_cast<float4*>(o4)[i] = reinterpret_cast<const float4*>(f.in[3])[i];
  }
}

__global__ __launch_bounds__(256) void copy(const float4* a, float4* b, size_t n4) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) b[i] = a[i];
}

#define CK(x)                                                                             \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) {                                                               \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                        \
      std::exit(1);                                                                       \
    }                                                                                     \
  } while (0)

template <class F>
float time_ms(F launch, int reps) {
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  for (int i = 0; i < 3; i++) launch();
  CK(hipEventRecord(a));
  for (int i = 0; i < reps; i++) launch();
  CK(hipEventRecord(b));
  CK(hipEventSynchronize(b));
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, a, b));
  CK(hipEventDestroy(a));
  CK(hipEventDestroy(b));
  return ms / reps;
}

}  // namespace

int main() {
  const size_t n = (size_t)kP * kH;
  Fields f{};
  for (int k = 0; k < kIn; k++) {
    float* p = nullptr;
    CK(hipMalloc(&p, n * sizeof(float)));
    CK(hipMemset(p, 0, n * sizeof(float)));
    f.in[k] = p;
  }
  unsigned char* cl = nullptr;
  CK(hipMalloc(&cl, n));
  CK(hipMemset(cl, 1, n));
  f.cls = cl;
  for (int k = 0; k < kOut; k++) CK(hipMalloc(&f.out[k], n * sizeof(float)));
  const double bytes = 29.0 * kW * kH;
  const int tiles_x = (kW + 59) / 60, reps = 20;
  std::printf("| pattern | shape | ms | GB/s (29 B x %d x %d) |\n|---|---|---|---|\n", kW, kH);
  for (int span : {32, 106}) {
    const int nsp = (kH + span - 1) / span;
    for (int wpb : {4, 8}) {
      const int groups = (tiles_x + wpb - 1) / wpb;
      const dim3 grid(nsp * groups), block(64 * wpb);
      float ms = time_ms([&] { hipLaunchKernelGGL(march<false>, grid, block, 0, 0, f, span, tiles_x, groups); }, reps);
      std::printf("| march | %d-row spans, %d waves/group | %.4f | %.0f |\n", span, wpb, ms, bytes / ms / 1e6);
      ms = time_ms([&] { hipLaunchKernelGGL(march<true>, grid, block, 0, 0, f, span, tiles_x, groups); }, reps);
      std::printf("| march/%d in step | %d-row spans | %.4f | %.0f |\n", wpb, span, ms, bytes / ms / 1e6);
    }
  }
  for (int span : {32, 160}) {
    const int nsp = (kH + span - 1) / span, groups = (tiles_x + 7) / 8;
    const float ms = time_ms([&] {
      hipLaunchKernelGGL(march_lds, dim3(nsp * groups), dim3(512), 0, 0, f, span, groups);
    }, reps);
    std::printf("| march/lds | %d-row spans, 8 waves/group, rows staged in LDS | %.4f | %.0f |\n", span, ms,
                bytes / ms / 1e6);
  }
  for (int span : {160}) {
    const int nsp = (kH + span - 1) / span, groups = (tiles_x + 7) / 8;
    const float ms = time_ms([&] {
      hipLaunchKernelGGL(march<true>, dim3(nsp * groups), dim3(512), 0, 0, f, span, tiles_x, groups);
    }, reps);
    std::printf("| march/8 in step | %d-row spans | %.4f | %.0f |\n", span, ms, bytes / ms / 1e6);
  }
  const size_t n4 = (size_t)kW * kH / 4;
  for (int blocks : {2048, 8192}) {
    const float ms = time_ms([&] { hipLaunchKernelGGL(flat, dim3(blocks), dim3(256), 0, 0, f, n4); }, reps);
    std::printf("| flat | %d blocks of 256 | %.4f | %.0f |\n", blocks, ms, 29.0 * 4 * n4 / ms / 1e6);
  }
  for (int blocks : {2048, 8192}) {
    const float ms = time_ms([&] { hipLaunchKernelGGL(flat_st<true>, dim3(blocks), dim3(256), 0, 0, f, n4); }, reps);
    std::printf("| flat/nt | %d blocks of 256 | %.4f | %.0f |\n", blocks, ms, 29.0 * 4 * n4 / ms / 1e6);
  }
  {
    const size_t nc = (size_t)kW * kH;
    float4 *a = nullptr, *b = nullptr;
    CK(hipMalloc(&a, nc * sizeof(float4)));
    CK(hipMalloc(&b, nc * sizeof(float4)));
    CK(hipMemset(a, 0, nc * sizeof(float4)));
    for (int blocks : {2048, 8192}) {
      float ms = time_ms([&] { hipLaunchKernelGGL(aos<false>, dim3(blocks), dim3(256), 0, 0, a, f.cls, b, nc); }, reps);
      std::printf("| aos | %d blocks of 256 | %.4f | %.0f (33 B per cell moved) |\n", blocks, ms, bytes / ms / 1e6);
      ms = time_ms([&] { hipLaunchKernelGGL(aos<true>, dim3(blocks), dim3(256), 0, 0, a, f.cls, b, nc); }, reps);
      std::printf("| aos/nt | %d blocks of 256 | %.4f | %.0f (33 B per cell moved) |\n", blocks, ms, bytes / ms / 1e6);
    }
    CK(hipFree(a));
    CK(hipFree(b));
  }
  const float ms = time_ms([&] {
    hipLaunchKernelGGL(copy, dim3(8192), dim3(256), 0, 0, reinterpret_cast<const float4*>(f.in[0]),
                       reinterpret_cast<float4*>(f.out[0]), n4);
  }, reps);
  std::printf("| copy | 1 read + 1 write, 8192 blocks | %.4f | %.0f (8 B per element) |\n", ms,
              8.0 * 4 * n4 / ms / 1e6);
  float* sink = nullptr;
  CK(hipMalloc(&sink, 256));
  float* o4 = nullptr;
  CK(hipMalloc(&o4, n * sizeof(float)));
  float m = time_ms([&] { hipLaunchKernelGGL(rd, dim3(8192), dim3(256), 0, 0, f, n4, sink); }, reps);
  std::printf("| read | 8192 blocks of 256 | %.4f | %.0f (17 B per cell moved: %.0f) |\n", m,
              bytes / m / 1e6, 17.0 * 4 * n4 / m / 1e6);
  m = time_ms([&] { hipLaunchKernelGGL(wr, dim3(8192), dim3(256), 0, 0, f, n4); }, reps);
  std::printf("| write | 8192 blocks of 256 | %.4f | %.0f (12 B per cell moved: %.0f) |\n", m,
              bytes / m / 1e6, 12.0 * 4 * n4 / m / 1e6);
  m = time_ms([&] { hipLaunchKernelGGL(copy4, dim3(8192), dim3(256), 0, 0, f, o4, n4); }, reps);
  std::printf("| copy4 | 4 reads + 4 writes, 8192 blocks | %.4f | %.0f (32 B per cell moved: %.0f) |\n", m,
              bytes / m / 1e6, 32.0 * 4 * n4 / m / 1e6);
  CK(hipFree(o4));
  CK(hipFree(sink));
  {
    // eight fields of fb bytes (a 2 MiB multiple) in one allocation; field k at k (fb + s)
    const size_t fb = (n * sizeof(float) + (2u << 20) - 1) / (2u << 20) * (2u << 20);
    const size_t smax = 1u << 20, total = 8 * (fb + smax);
    char* big = nullptr;
    CK(hipMalloc(&big, total));
    CK(hipMemset(big, 0, total));
    for (size_t sg : {(size_t)0, (size_t)4096, (size_t)65536, smax}) {
      Fields h{};
      size_t off = 0;
      for (int k = 0; k < kIn; k++, off += fb + sg) h.in[k] = reinterpret_cast<const float*>(big + off);
      h.cls = reinterpret_cast<const unsigned char*>(big + off);
      off += fb + sg;
      for (int k = 0; k < kOut; k++, off += fb + sg) h.out[k] = reinterpret_cast<float*>(big + off);
      if (off > total) return 1;
      m = time_ms([&] { hipLaunchKernelGGL(flat, dim3(8192), dim3(256), 0, 0, h, n4); }, reps);
      std::printf("| stagger | field bases %zu B apart beyond their size | %.4f | %.0f |\n", sg, m,
                  29.0 * 4 * n4 / m / 1e6);
    }
    CK(hipFree(big));
  }
  {
    // 64-column segments over the pitch: the SoA fields above hold kP * kH elements each
    static_assert(kP % 64 == 0, "segments cover the pitch");
    const size_t nseg = (size_t)kP / 64 * kH;
    float *hqq = nullptr, *out = nullptr;
    CK(hipMalloc(&hqq, 3 * n * sizeof(float)));
    CK(hipMalloc(&out, 3 * n * sizeof(float)));
    CK(hipMemset(hqq, 0, 3 * n * sizeof(float)));
    for (int blocks : {2048, 8192}) {
      m = time_ms([&] { hipLaunchKernelGGL(seg_soa<false>, dim3(blocks), dim3(256), 0, 0, f, nseg); }, reps);
      std::printf("| seg/soa | %d blocks of 256, 8 streams | %.4f | %.0f |\n", blocks, m, bytes / m / 1e6);
      m = time_ms([&] { hipLaunchKernelGGL(seg_soa<true>, dim3(blocks), dim3(256), 0, 0, f, nseg); }, reps);
      std::printf("| seg/soa/nt | %d blocks of 256, 8 streams | %.4f | %.0f |\n", blocks, m, bytes / m / 1e6);
      m = time_ms([&] {
        hipLaunchKernelGGL(seg_aosoa<false>, dim3(blocks), dim3(256), 0, 0, hqq, f.in[3], f.cls, out, nseg);
      }, reps);
      std::printf("| seg/aosoa | %d blocks of 256, 4 streams | %.4f | %.0f |\n", blocks, m, bytes / m / 1e6);
      m = time_ms([&] {
        hipLaunchKernelGGL(seg_aosoa<true>, dim3(blocks), dim3(256), 0, 0, hqq, f.in[3], f.cls, out, nseg);
      }, reps);
      std::printf("| seg/aosoa/nt | %d blocks of 256, 4 streams | %.4f | %.0f |\n", blocks, m, bytes / m / 1e6);
    }
    CK(hipFree(hqq));
    CK(hipFree(out));
  }
  CK(hipDeviceSynchronize());
  return 0;
}
