// Per-step control logic shared by the precompiled kernels and the run-time compiled
// boundary kernel (this file is also handed to hipRTC as a header).
#pragma once

// controlTimestep (src/main.cpp:25-43), from the CFL minimum of the last step. The
// reference scans a per-tile table whose entries start at 1; its minimum equals the
// running minimum over wet stored cells (capped at 1) kept in cfl_prev.
__device__ __forceinline__ float fip_control_dt(float dt, float cfl_prev) {
  float min_cfl = 42.0f;
  if (cfl_prev < min_cfl) min_cfl = cfl_prev;
  const float target = min_cfl * .5f;
  const float dc = 0.001f;
  if (dt < target && target - dt > dc)
    dt *= 1.001f;
  else if (dt > target && dt - target > dc)
    dt *= 0.99f;
  return dt;
}
__device__ __forceinline__ void fip_control(fip_ctl* c) { c->dt = fip_control_dt(c->dt, c->cfl_prev); }

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

// sample() (src/main.cpp:53-83), run at the start of a simulate() call (main.cpp:108-112):
// when duration >= Sampling::timer the timer advances by the interval and one row is
// logged. fip_sample_begin runs on one thread, updates the timer and reserves the row
// (returns -1 when nothing is logged); fip_sample_point then fills one point of it. Type
// z adds the bed to the depth.
__device__ __forceinline__ int fip_sample_begin(fip_ctl* c, double t, double timer, int interval,
                                                int nrows) {
  if (!(t >= timer && interval > 0)) return -1;
  c->sample_timer = timer + (double)interval;
  if (nrows >= FIP_SAMPLE_LOG_ROWS) {
    c->sample_overflow = 1;
    return -1;
  }
  c->sample_t[nrows] = t;
  c->nsample_rows = nrows + 1;
  return nrows;
}

__device__ __forceinline__ void fip_sample_point(const fip_grid& g, int cur, int row, int p) {
  fip_ctl* c = g.ctl;
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
