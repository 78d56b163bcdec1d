// Active-tile scheduling for the masked stencil, shared by the precompiled library and
// the run-time compiled boundary kernel (this file is also handed to hipRTC).
//
// Tile (tx, ty) covers columns [60 tx, 60 tx + 60) and rows [R ty, R ty + R). Its 64-bit
// mask word has bit k set when a stored depth in row R ty + k ended wet (> 1e-6). The
// rows to update next are the set bits of the dilation used by eval_mask
// (src/kernels.cu:129-154), at this geometry and with neighbours beyond the grid dry:
//   need = own | own<<1 | own>>1 | west | east | (north & 1) << (R-1) | (south >> (R-1)) & 1
// A tile with need == 0 is not listed (never launched). A listed tile stores rows
// [ctz(need), 63 - clz(need)] of its own rows.
#pragma once

__device__ __forceinline__ unsigned long long fip_valid_rows(int nrows) {
  return nrows >= 64 ? ~0ull : ((1ull << nrows) - 1ull);
}

__device__ __forceinline__ unsigned long long fip_tile_need(const fip_grid& g, int t) {
  const int TX = g.tiles_x, TY = g.tiles_y, R = g.tile_rows;
  const int tx = t % TX, ty = t / TX;
  const int nrows = (g.H - ty * R) < R ? (g.H - ty * R) : R;
  const unsigned long long* m = g.mask;
  unsigned long long o = m[t];
  unsigned long long w = tx > 0 ? m[t - 1] : 0ull;
  unsigned long long e = tx < TX - 1 ? m[t + 1] : 0ull;
  unsigned long long n = ty < TY - 1 ? m[t + TX] : 0ull;
  unsigned long long s = ty > 0 ? m[t - TX] : 0ull;
  unsigned long long need =
      o | (o << 1) | (o >> 1) | w | e | ((n & 1ull) << (R - 1)) | ((s >> (R - 1)) & 1ull);
  return need & fip_valid_rows(nrows);
}

// Positions in a list of length *count for the threads of a work-group that pass act:
// one atomic per work-group and call (one per wavefront put hundreds of atomics on the
// one counter each step, and they serialise). Every thread of the work-group must make
// the call, the same number of times (it contains barriers). Returns this thread's
// position when act; positions follow the thread order within the work-group.
__device__ __forceinline__ unsigned fip_block_append(bool act, unsigned* count) {
  __shared__ unsigned s_cnt[16];
  __shared__ unsigned s_base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  const unsigned long long bal = __ballot(act);
  if (lane == 0) s_cnt[wave] = (unsigned)__popcll(bal);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned tot = 0u;
    for (int w = 0; w < nw; w++) tot += s_cnt[w];
    s_base = tot ? atomicAdd(count, tot) : 0u;
  }
  __syncthreads();
  unsigned pos = s_base;
  for (int w = 0; w < wave; w++) pos += s_cnt[w];
  pos += (unsigned)__popcll(bal & ((1ull << lane) - 1ull));
  __syncthreads();  // s_cnt and s_base are rewritten by the next call
  return pos;
}

// Appends the active tiles to list[q]; called by every thread of `nblocks` work-groups of
// at most 1024 threads (every thread of a work-group runs the same iterations).
__device__ __forceinline__ void fip_schedule(const fip_grid& g, int q, int block, int nblocks) {
  const int ntiles = g.tiles_x * g.tiles_y;
  const int stride = nblocks * blockDim.x;
  for (int base = block * blockDim.x; base < ntiles; base += stride) {
    const int t = base + threadIdx.x;
    const unsigned long long need = t < ntiles ? fip_tile_need(g, t) : 0ull;
    const bool act = need != 0ull;
    const unsigned pos = fip_block_append(act, &g.ctl->list_count[q]);
    if (act && pos < (unsigned)ntiles) {
      const unsigned lo = (unsigned)__builtin_ctzll(need);
      const unsigned hi = 63u - (unsigned)__builtin_clzll(need);
      g.list[q][pos] = ((unsigned)t << 12) | (lo << 6) | hi;
    }
  }
}
