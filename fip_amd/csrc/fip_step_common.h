// Per-step control logic shared by the precompiled kernels and the run-time compiled
// boundary kernel (this file is also handed to hipRTC as a header).
#pragma once

// controlTimestep (src/main.cpp:25-43), from the CFL minimum of the last step. The
// reference scans a per-tile table whose entries start at 1; its minimum equals the
// running minimum over wet stored cells (capped at 1) kept in cfl_prev.
__device__ __forceinline__ void fip_control(fip_ctl* c) {
  float min_cfl = 42.0f;
  if (c->cfl_prev < min_cfl) min_cfl = c->cfl_prev;
  const float target = min_cfl * .5f;
  const float dc = 0.001f;
  float dt = c->dt;
  if (dt < target && target - dt > dc)
    dt *= 1.001f;
  else if (dt > target && dt - target > dc)
    dt *= 0.99f;
  c->dt = dt;
}

// TimeSeries::interp (src/timeseries.cpp:14-26): std::map::lower_bound, clamped at both
// ends, then f * upper + (1 - f) * lower.
__device__ __forceinline__ float fip_interp(const float* t, const float* v, int n, float time) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (t[mid] < time)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo == 0) return v[0];
  if (lo == n) return v[n - 1];
  const int l = lo - 1;
  const float dt = t[lo] - t[l];
  const float f = (time - t[l]) / dt;
  return f * v[lo] + (1.f - f) * v[l];
}

// sample() (src/main.cpp:53-83) at the start of a simulate() call: logs (duration,
// values) when duration >= Sampling::timer. Type z adds the bed to the depth.
__device__ __forceinline__ void fip_sample(const fip_grid& g, int cur) {
  fip_ctl* c = g.ctl;
  if (!(c->duration >= c->sample_timer && c->sampling_interval > 0)) return;
  c->sample_timer = c->sample_timer + (double)c->sampling_interval;
  const int row = c->nsample_rows;
  if (row >= FIP_SAMPLE_LOG_ROWS) {
    c->sample_overflow = 1;
    return;
  }
  c->sample_t[row] = c->duration;
  for (int p = 0; p < c->nsample_points; p++) {
    const long i = (long)c->sp_y[p] * g.P + c->sp_x[p];
    float v;
    switch (c->sp_type[p]) {
      case 0: v = g.h[cur][i]; v += g.bed[i]; break;
      case 1: v = g.h[cur][i]; break;
      case 2: v = g.qx[cur][i]; break;
      default: v = g.qy[cur][i]; break;
    }
    c->sample_v[row * FIP_MAX_SAMPLE_POINTS + p] = v;
  }
  c->nsample_rows = row + 1;
}
