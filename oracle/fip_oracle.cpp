// CPU oracle for the fip shallow-water step. TEST INFRASTRUCTURE ONLY (see fip_oracle.h).
//
// Three restatements of the stencil, which must agree where their semantics meet:
//   * step_block   - a field-pass form of src/kernels.cu:35-127,351-406: every derived
//                    quantity (ux, uy, fx, fy, gx, gy, new qx/qy, new h, CFL) is formed
//                    once per cell from the old state with explicit neighbour indices.
//                    DENSE mode runs it on the whole grid; TILED mode runs it on the
//                    active rows of each tile of a given geometry (the HIP geometry).
//   * step_warp_ref - a literal transcription of rmg_kernel<50,false>
//                    (src/kernels.cu:264-438) with warp lanes emulated, including the
//                    per-warp mask (eval_mask/store_mask, :129-171) and CFL table
//                    (store_courant, :173-192).
// Arithmetic follows the C++ expression trees of the reference exactly (left-to-right
// association, no contraction: build with -ffp-contract=off -fno-fast-math).
#include "fip_oracle.h"
#include "../fip_amd/csrc/fip_detmath.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <class R> struct Num;
template <> struct Num<float> {
  static float max(float a, float b) { return fip_fmaxf(a, b); }
  static float min(float a, float b) { return fip_fminf(a, b); }
  static float pow233(float x) { return fip_pow233(x); }
  static float sqrt(float x) { return std::sqrt(x); }
  static float fabs(float x) { return std::fabs(x); }
};
template <> struct Num<double> {
  static double max(double a, double b) {
    if (a != a) return b;
    if (b != b) return a;
    if (a > b) return a;
    if (b > a) return b;
    return std::signbit(a) ? b : a;
  }
  static double min(double a, double b) {
    if (a != a) return b;
    if (b != b) return a;
    if (a < b) return a;
    if (b < a) return b;
    return std::signbit(a) ? a : b;
  }
  static double pow233(double x) { return std::pow(x, (double)2.33f); }
  static double sqrt(double x) { return std::sqrt(x); }
  static double fabs(double x) { return std::fabs(x); }
};

// Constants of src/kernels.cu:23-24 and src/simdata.h:48, widened in binary64.
template <class R> struct K {
  static constexpr R G = R(9.81f);
  static constexpr R DH = R(0.01f);
  static constexpr R EPS = R(10e-7f);
};

// fp32div_wrapper (src/kernels.cu:30-33): +0 for a zero dividend of either sign.
template <class R> inline R divw(R a, R b) { return (a == R(0)) ? R(0) : (a / b); }

// compute_rmg_qx (src/kernels.cu:75-100).
template <class R>
inline R rmg_qx(R fx_east, R fx_i, R gx_i, R gx_south, R dt, R dx, R dt_inverse, R dx_inverse,
                R kSt, R h_i, R h_east, R sohle_east, R sohle_i, R qx_old_i, R qy_old_i,
                R qy_old_east, R qy_old_south, R qy_old_southeast) {
  typedef Num<R> N;
  const R G = K<R>::G, DH = K<R>::DH;
  R dfx = (fx_east - fx_i) * dx_inverse;
  R dgx = (gx_i - gx_south) * dx_inverse;
  R ah_qx = (h_i + h_east) * R(.5f);
  ah_qx = N::max(R(0.f), ah_qx);
  R dhg = ah_qx * G * (-sohle_east + sohle_i) * dx_inverse;
  R dhx = (R(0.5f) * G * h_east * h_east - R(0.5f) * G * h_i * h_i) * dx_inverse;
  ah_qx = N::max(ah_qx, DH);
  const R ah_qx_inverted = R(1.f) / ah_qx;
  const R ux = qx_old_i;
  const R uy = (qy_old_i + qy_old_east + qy_old_south + qy_old_southeast) * R(0.25f);
  R reib = ((ux == R(0.f)) && (uy == R(0.f)))
               ? R(0.f)
               : (dt * G / (kSt * kSt) * N::sqrt(ux * ux + uy * uy) * N::pow233(ah_qx_inverted));
  R qx_new = divw<R>(qx_old_i + dt * (-dfx - dgx - dhx + dhg), (R(1.f) + reib));
  qx_new = N::min(qx_new, h_i * dt_inverse * dx * R(0.25f));
  qx_new = N::max(qx_new, -h_east * dt_inverse * dx * R(0.25f));
  return qx_new;
}

// compute_rmg_qy (src/kernels.cu:102-127).
template <class R>
inline R rmg_qy(R fy_i, R fy_west, R gy_i, R gy_north, R dt, R dx, R dt_inverse, R dx_inverse,
                R kSt, R h_i, R h_north, R sohle_north, R sohle_i, R qy_old_i, R qx_old_i,
                R qx_old_west, R qx_old_north, R qx_old_northwest) {
  typedef Num<R> N;
  const R G = K<R>::G, DH = K<R>::DH;
  R dgy = (gy_north - gy_i) * dx_inverse;
  R dfy = (fy_i - fy_west) * dx_inverse;
  R ah_qy = (h_i + h_north) * R(.5f);
  ah_qy = N::max(R(0.f), ah_qy);
  R dhg = ah_qy * G * (sohle_i - sohle_north) * dx_inverse;
  R dhy = (R(0.5f) * G * h_north * h_north - R(0.5f) * G * h_i * h_i) * dx_inverse;
  ah_qy = N::max(ah_qy, DH);
  R ah_qy_inverted = R(1.f) / ah_qy;
  R ux = (qx_old_i + qx_old_north + qx_old_west + qx_old_northwest) * R(0.25f);
  R uy = qy_old_i;
  R reib = ((ux == R(0.f)) && (uy == R(0.f)))
               ? R(0.f)
               : (dt * G / (kSt * kSt) * N::sqrt(ux * ux + uy * uy) * N::pow233(ah_qy_inverted));
  R qy_new = divw<R>(qy_old_i + dt * (-dgy - dfy - dhy + dhg), (R(1) + reib));
  qy_new = N::min(qy_new, h_i * dt_inverse * dx * R(0.25f));
  qy_new = N::max(qy_new, -h_north * dt_inverse * dx * R(0.25f));
  return qy_new;
}

struct Series {
  std::vector<float> t, v;
  // TimeSeries::interp (src/timeseries.cpp:14-26) over a std::map<float,float>.
  float interp(float time) const {
    size_t n = t.size();
    size_t up = std::lower_bound(t.begin(), t.end(), time) - t.begin();
    if (up == 0) return v[0];
    if (up == n) return v[n - 1];
    size_t lo = up - 1;
    float dt = t[up] - t[lo];
    float f = (time - t[lo]) / dt;
    return f * v[up] + (1.f - f) * v[lo];
  }
};

struct BC {
  int type, side, from, to;
  float z, q;
  bool series;
  Series zs, qs;
  bool has_zs, has_qs;
};

template <class R> struct St {
  int W, H, P;  // pitch in elements
  R dx, dt;
  R no_data;
  R kst;
  bool has_kvar;
  int mode, tw, tr, warp_mask, quirks, variable_dt, staggered, threads;
  int open_pitch = 0;
  std::vector<R> bed, kv, hmax;
  std::vector<R> h[2], qx[2], qy[2];
  std::vector<float> bed_host;  // the host copy of the bed (h_sohle), float
  int cur = 0;
  double duration = 0;
  int iteration = 0;
  // CFL: DENSE/TILED keep the running minimum; WARP_REF keeps the per-warp table.
  R cfl_acc = R(1.f);
  std::vector<R> cfl_table;
  std::vector<uint64_t> mask;  // TILED: one word per tile; WARP_REF: per warp
  int tiles_x = 0, tiles_y = 0;
  long long cells_updated = 0, tiles_computed = 0;
  std::vector<BC> bcs;

  size_t idx(int i, int j) const { return (size_t)j * P + i; }
};

// ------------------------------------------------------------------------------------------
// Field-pass restatement on a block of stored cells [i0,i1) x [j0,j1).
// Returns the CFL minimum over stored wet cells (1 if none) and ORs the wet rows
// (relative to rowbase) into *rows.
template <class R>
R step_block(St<R>& s, int i0, int i1, int j0, int j1, int rowbase, uint64_t* rows,
             long long* ncells) {
  typedef Num<R> N;
  const R G = K<R>::G, DH = K<R>::DH, EPS = K<R>::EPS;
  const int W = s.W, H = s.H;
  const R* h = s.h[s.cur].data();
  const R* qx = s.qx[s.cur].data();
  const R* qy = s.qy[s.cur].data();
  const R* b = s.bed.data();
  R* h2o = s.h[s.cur ^ 1].data();
  R* qx2o = s.qx[s.cur ^ 1].data();
  R* qy2o = s.qy[s.cur ^ 1].data();
  const R dt = s.dt, dx = s.dx;
  const R dx_inverse = R(1.f) / dx;
  const R dt_inverse = R(1.f) / dt;

  // box with a 2-cell halo; outside the grid every old value reads as 0
  const int bi0 = i0 - 2, bj0 = j0 - 2;
  const int bw = (i1 - i0) + 4, bh = (j1 - j0) + 4;
  auto in = [&](int i, int j) { return i >= 0 && i < W && j >= 0 && j < H; };
  auto Hh = [&](int i, int j) { return in(i, j) ? h[s.idx(i, j)] : R(0); };
  auto QX = [&](int i, int j) { return in(i, j) ? qx[s.idx(i, j)] : R(0); };
  auto QY = [&](int i, int j) { return in(i, j) ? qy[s.idx(i, j)] : R(0); };
  auto B = [&](int i, int j) { return in(i, j) ? b[s.idx(i, j)] : R(0); };
  auto KS = [&](int i, int j) {
    return s.has_kvar ? (in(i, j) ? s.kv[s.idx(i, j)] : R(0)) : s.kst;
  };
  std::vector<R> UX(bw * bh), UY(bw * bh), FX(bw * bh), FY(bw * bh), GX(bw * bh), GY(bw * bh),
      QX2(bw * bh), QY2(bw * bh);
  auto L = [&](int i, int j) { return (size_t)(j - bj0) * bw + (i - bi0); };
  auto inb = [&](int i, int j) { return i >= bi0 && i < bi0 + bw && j >= bj0 && j < bj0 + bh; };
  auto get = [&](const std::vector<R>& a, int i, int j) { return inb(i, j) ? a[L(i, j)] : R(0); };

  for (int j = bj0; j < bj0 + bh; j++)
    for (int i = bi0; i < bi0 + bw; i++) {
      // ux on the east face (i+1/2), uy on the north face (j+1/2): kernels.cu:351-354
      UX[L(i, j)] = divw<R>(QX(i, j), N::max(N::max(DH, Hh(i, j)), Hh(i + 1, j)));
      UY[L(i, j)] = divw<R>(QY(i, j), N::max(N::max(DH, Hh(i, j)), Hh(i, j + 1)));
    }
  for (int j = bj0; j < bj0 + bh; j++)
    for (int i = bi0; i < bi0 + bw; i++) {
      // compute_rmg_fx (kernels.cu:56-63)
      R uw = get(UX, i - 1, j), ui = get(UX, i, j);
      FX[L(i, j)] = (uw + ui > 0) ? QX(i - 1, j) * uw : QX(i, j) * ui;
      // compute_rmg_fy (kernels.cu:45-54)
      R un = get(UX, i, j + 1);
      R uxh = (ui + un) * R(.5f);
      FY[L(i, j)] = (ui + un > 0) ? QY(i, j) * uxh : QY(i + 1, j) * uxh;
      // compute_rmg_gx (kernels.cu:35-43)
      R vi = get(UY, i, j), ve = get(UY, i + 1, j);
      R uyh = (vi + ve) * R(.5f);
      GX[L(i, j)] = (vi + ve > 0) ? QX(i, j) * uyh : QX(i, j + 1) * uyh;
      // compute_rmg_gy (kernels.cu:65-73), the flux on the south face of cell j
      R vs = get(UY, i, j - 1);
      GY[L(i, j)] = (vi + vs > 0) ? QY(i, j - 1) * vs : QY(i, j) * vi;
    }
  for (int j = bj0; j < bj0 + bh; j++)
    for (int i = bi0; i < bi0 + bw; i++) {
      R qxn = rmg_qx<R>(get(FX, i + 1, j), get(FX, i, j), get(GX, i, j), get(GX, i, j - 1), dt,
                        dx, dt_inverse, dx_inverse, KS(i, j), Hh(i, j), Hh(i + 1, j),
                        B(i + 1, j), B(i, j), QX(i, j), QY(i, j), QY(i + 1, j), QY(i, j - 1),
                        QY(i + 1, j - 1));
      R qyn = rmg_qy<R>(get(FY, i, j), get(FY, i - 1, j), get(GY, i, j), get(GY, i, j + 1), dt,
                        dx, dt_inverse, dx_inverse, KS(i, j), Hh(i, j), Hh(i, j + 1),
                        B(i, j + 1), B(i, j), QY(i, j), QX(i, j), QX(i - 1, j), QX(i, j + 1),
                        QX(i - 1, j + 1));
      // faces held for flow boundary conditions (kernels.cu:373-376)
      if (i == 0 || i == W - 2) qxn = QX(i, j);
      if (j == 0 || j == H - 2) qyn = QY(i, j);
      QX2[L(i, j)] = qxn;
      QY2[L(i, j)] = qyn;
    }
  R cfl = R(1.f);
  long long n = 0;
  for (int j = j0; j < j1; j++)
    for (int i = i0; i < i1; i++) {
      if (!(B(i, j) != s.no_data)) continue;  // kernels.cu:383
      const size_t g = s.idx(i, j);
      const bool sh = i >= 1 && i <= W - 2 && j >= 1 && j <= H - 2;
      const bool sqx = i >= 1 && i <= W - 3 && j >= 2 && j <= H - 3;
      const bool sqy = i >= 2 && i <= W - 3 && j >= 1 && j <= H - 3;
      if (sh) {
        // continuity (kernels.cu:381)
        R h2 = Hh(i, j) - dt * dx_inverse *
                              ((get(QX2, i, j) - get(QX2, i - 1, j)) +
                               (get(QY2, i, j) - get(QY2, i, j - 1)));
        h2o[g] = h2;
        n++;
        if (h2 > EPS) {
          R umax = N::max(N::fabs(get(UY, i, j)),
                          N::max(N::fabs(get(UY, i, j - 1)),
                                 N::max(N::fabs(get(UX, i, j)), N::fabs(get(UX, i - 1, j)))));
          R dt_cfl = dx / (N::sqrt(G * h2) + umax);
          cfl = N::min(cfl, dt_cfl);
          if (rows) *rows |= 1ull << (j - rowbase);
        }
      }
      if (sqx) qx2o[g] = get(QX2, i, j);
      if (sqy) qy2o[g] = get(QY2, i, j);
    }
  if (ncells) *ncells += n;
  return cfl;
}

template <class R> void stencil_dense(St<R>& s) {
  const int band = 64;
  const int nb = (s.H + band - 1) / band;
  R cfl = R(1.f);
  long long cells = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : cells) num_threads(s.threads)
  for (int t = 0; t < nb; t++) {
    long long c = 0;
    R m = step_block<R>(s, 0, s.W, t * band, std::min(s.H, (t + 1) * band), 0, nullptr, &c);
    cells += c;
#pragma omp critical
    cfl = Num<R>::min(cfl, m);
  }
  s.cfl_acc = cfl;
  s.cells_updated = cells;
  s.tiles_computed = nb;
}

inline uint64_t valid_rows(int nrows) {
  return nrows >= 64 ? ~0ull : ((1ull << nrows) - 1);
}

// Wet/dry tile skipping at geometry (tw columns x tr rows per tile). Each tile owns
// columns [tw*tx, tw*tx+tw) and rows [tr*ty, tr*ty+tr). Its mask word has bit k set when
// a stored depth in row tr*ty+k ended wet (> EPS). Before a step the rows to update are
// the set bits of
//   need = own | own<<1 | own>>1 | west | east | (north&1)<<(tr-1) | (south>>(tr-1))&1,
// (the dilation of eval_mask, src/kernels.cu:129-154, at this geometry; neighbours
// outside the grid count as dry). A tile with need == 0 is skipped and its output cells
// keep their previous contents. Otherwise rows [ctz(need), 63-clz(need)] of the tile are
// stored exactly as in dense mode.
template <class R> void stencil_tiled(St<R>& s) {
  const int TX = s.tiles_x, TY = s.tiles_y, tw = s.tw, tr = s.tr;
  std::vector<uint64_t> own = s.mask;
  R cfl = R(1.f);
  long long cells = 0, tiles = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : cells, tiles) num_threads(s.threads)
  for (int t = 0; t < TX * TY; t++) {
    const int tx = t % TX, ty = t / TX;
    const int nrows = std::min(tr, s.H - ty * tr);
    uint64_t o = own[t];
    uint64_t w = tx > 0 ? own[t - 1] : 0;
    uint64_t e = tx < TX - 1 ? own[t + 1] : 0;
    uint64_t n = ty < TY - 1 ? own[t + TX] : 0;
    uint64_t so = ty > 0 ? own[t - TX] : 0;
    uint64_t need = o | (o << 1) | (o >> 1) | w | e | ((n & 1ull) << (tr - 1)) |
                    ((so >> (tr - 1)) & 1ull);
    need &= valid_rows(nrows);
    if (need == 0) continue;
    const int lo = __builtin_ctzll(need), hi = 63 - __builtin_clzll(need);
    uint64_t rows = 0;
    long long c = 0;
    R m = step_block<R>(s, tx * tw, std::min(s.W, tx * tw + tw), ty * tr + lo, ty * tr + hi + 1,
                        ty * tr, &rows, &c);
    s.mask[t] = rows;
    cells += c;
    tiles++;
#pragma omp critical
    cfl = Num<R>::min(cfl, m);
  }
  s.cfl_acc = cfl;
  s.cells_updated = cells;
  s.tiles_computed = tiles;
}

// ------------------------------------------------------------------------------------------
// Literal transcription of rmg_kernel<BLOCKROWS,false> with emulated warp lanes.
// Shuffles follow CUDA: __shfl_down(x,1) reads lane+1 (own value at the last lane),
// __shfl_up(x,1) reads lane-1 (own value at lane 0). Lanes that exited at col >= W read
// as 0. Mask reads are taken from the masks as they were before the launch (the
// reference reads neighbours while they are being rewritten; the oracle fixes the
// order), and reads past the written tiles return all ones (a never-written entry).
template <class R> void stencil_warp_ref(St<R>& s) {
  typedef Num<R> N;
  const R G = K<R>::G, DH = K<R>::DH, EPS = K<R>::EPS;
  const int LANES = s.tw, BLOCKROWS = s.tr, STENCIL = 2;
  const int W = s.W, H = s.H, pitch = s.P;
  const int wptb = 4;
  const int cols = LANES - 2 * STENCIL;
  const int gx = W / (cols * wptb) + 1, gy = H / BLOCKROWS + 1;
  const int warpsPerRow = gx * wptb;
  const int ntiles = warpsPerRow * gy;
  const std::vector<uint64_t> mold = s.mask;
  auto mread = [&](long idx) -> uint64_t {
    if (idx < 0 || idx >= (long)ntiles) return ~0ull;
    return mold[idx];
  };
  const R* h1 = s.h[s.cur].data();
  const R* qx1 = s.qx[s.cur].data();
  const R* qy1 = s.qy[s.cur].data();
  const R* sohle = s.bed.data();
  R* h2 = s.h[s.cur ^ 1].data();
  R* qx2 = s.qx[s.cur ^ 1].data();
  R* qy2 = s.qy[s.cur ^ 1].data();
  const R dt = s.dt, dx = s.dx;
  const R dx_inverse = R(1.f) / dx;
  const R dt_inverse = R(1.f) / dt;
  long long cells = 0, tiles = 0;

  // With the literal store predicates (quirks) two tiles can store the same cell; tiles
  // then run in order, so the later tile's value is kept, deterministically.
#pragma omp parallel for schedule(dynamic) reduction(+ : cells, tiles) num_threads(s.threads) if (!s.quirks)
  for (int t = 0; t < ntiles; t++) {
    const int by = t / warpsPerRow, wx = t % warpsPerRow;
    std::vector<int> col(LANES);
    std::vector<char> act(LANES);
    int row0 = by * BLOCKROWS;
    bool any = false;
    for (int l = 0; l < LANES; l++) {
      col[l] = wx * cols + l;
      act[l] = !(col[l] >= W || row0 >= H);
      any |= act[l];
    }
    if (!any) continue;
    int firstRow = 0, lastRow = BLOCKROWS + STENCIL;
    if (s.warp_mask) {
      // eval_mask (kernels.cu:129-154), evaluated by lane 0
      const int c0 = col[0];
      long maskIdx = (long)by * warpsPerRow + wx;
      uint64_t m = mread(maskIdx);
      uint64_t m_west = (c0 == 0) ? 0 : mread(maskIdx - 1);
      uint64_t m_east = (c0 == W - 1) ? 0 : mread(maskIdx + 1);
      uint64_t m_north = (m >> 1);
      uint64_t m_south = (m << 1);
      m_north |= (row0 == H - 1) ? 0
                                 : ((mread(maskIdx + warpsPerRow) & (uint64_t)(1 << STENCIL))
                                    << (BLOCKROWS - 1));
      m_south |= (row0 == 0) ? 0 : (mread(maskIdx - warpsPerRow) >> (BLOCKROWS - 1));
      m = m | m_south | m_north | m_west | m_east;
      if (m == 0) continue;
      firstRow = std::max(0, __builtin_ffsll(m) - STENCIL - 1);
      lastRow = std::min(BLOCKROWS + STENCIL, 64 - __builtin_clzll(m));
    }
    tiles++;
    int row = row0 + firstRow;
    const int totalrows = std::min(lastRow - firstRow, H - row);
    int gi = row * pitch;  // + col per lane

    std::vector<R> h_i(LANES), qx1_i(LANES), qy1_i(LANES), sohle_north(LANES), h_north(LANES),
        h_northnorth(LANES), qy1_north(LANES), qx1_north(LANES), h_east(LANES),
        h_northeast(LANES), qx1_northwest(LANES), qy1_east(LANES), rb(LANES), sohle_i(LANES),
        sohle_east(LANES), qy1_south(LANES), qx1_west(LANES), qy1_southeast(LANES),
        gx_south(LANES), gy_i(LANES), qy2_south(LANES), uy_south(LANES), dtc(LANES);
    std::vector<uint64_t> mnew(LANES, 0);
    auto ld = [&](const R* a, int l, int off) -> R { return act[l] ? a[gi + col[l] + off] : R(0); };
    auto down = [&](const std::vector<R>& v, int l) -> R {
      if (l + 1 >= LANES) return act[l] ? v[l] : R(0);
      return act[l + 1] ? v[l + 1] : R(0);
    };
    auto up = [&](const std::vector<R>& v, int l) -> R {
      if (l == 0) return act[l] ? v[l] : R(0);
      return act[l - 1] ? v[l - 1] : R(0);
    };
    for (int l = 0; l < LANES; l++) {
      h_i[l] = ld(h1, l, 0);
      qx1_i[l] = ld(qx1, l, 0);
      qy1_i[l] = ld(qy1, l, 0);
      sohle_north[l] = ld(sohle, l, pitch);
      h_north[l] = ld(h1, l, pitch);
      h_northnorth[l] = ld(h1, l, 2 * pitch);
      qy1_north[l] = ld(qy1, l, pitch);
      qx1_north[l] = ld(qx1, l, pitch);
      dtc[l] = R(1.f);
    }
    for (int l = 0; l < LANES; l++) {
      h_east[l] = down(h_i, l);
      h_northeast[l] = down(h_north, l);
      qx1_northwest[l] = up(qx1_north, l);
      qy1_east[l] = down(qy1_i, l);
      rb[l] = sohle_i[l] = sohle_east[l] = qy1_south[l] = qx1_west[l] = qy1_southeast[l] =
          gx_south[l] = gy_i[l] = qy2_south[l] = uy_south[l] = R(0);
    }
    std::vector<R> rb_n(LANES), so_nn(LANES), h_nnn(LANES), qx_nn(LANES), qy_nn(LANES);
    std::vector<R> ux_i(LANES), uy_i(LANES), fx_i(LANES), fy_i(LANES), gx_i(LANES),
        gy_north(LANES), qx2_i(LANES), qy2_i(LANES), ux_w(LANES), uy_e(LANES), ux_iA(LANES);
    for (int i = 0; i < totalrows; row++, gi += pitch, i++) {
      for (int l = 0; l < LANES; l++) {
        mnew[l] >>= 1;
        rb_n[l] = so_nn[l] = h_nnn[l] = qx_nn[l] = qy_nn[l] = R(0.f);
        if (i != totalrows - 1) {
          rb_n[l] = s.has_kvar ? ld(s.kv.data(), l, pitch) : s.kst;
          so_nn[l] = ld(sohle, l, 2 * pitch);
          h_nnn[l] = ld(h1, l, 3 * pitch);
          qx_nn[l] = ld(qx1, l, 2 * pitch);
          qy_nn[l] = ld(qy1, l, 2 * pitch);
        }
        R ah_qx = N::max(N::max(DH, h_i[l]), h_east[l]);
        R ah_qy = N::max(N::max(DH, h_i[l]), h_north[l]);
        ux_i[l] = divw<R>(qx1_i[l], ah_qx);
        uy_i[l] = divw<R>(qy1_i[l], ah_qy);
      }
      for (int l = 0; l < LANES; l++) {
        ux_w[l] = up(ux_i, l);
        uy_e[l] = down(uy_i, l);
      }
      for (int l = 0; l < LANES; l++) {
        // compute_rmg_fx
        fx_i[l] = (ux_w[l] + ux_i[l] > 0) ? qx1_west[l] * ux_w[l] : qx1_i[l] * ux_i[l];
        // compute_rmg_fy
        R ah_qx_north = N::max(N::max(DH, h_north[l]), h_northeast[l]);
        R ux_north = divw<R>(qx1_north[l], ah_qx_north);
        R ux_h = (ux_i[l] + ux_north) * R(.5f);
        fy_i[l] = (ux_i[l] + ux_north > 0) ? qy1_i[l] * ux_h : qy1_east[l] * ux_h;
        // compute_rmg_gx
        R uy_h = (uy_i[l] + uy_e[l]) * R(.5f);
        gx_i[l] = (uy_i[l] + uy_e[l] > 0) ? qx1_i[l] * uy_h : qx1_north[l] * uy_h;
        // compute_rmg_gy(h_north, h_northnorth, h_i, qy1_north, qy1_i)
        R a1 = N::max(N::max(DH, h_north[l]), h_northnorth[l]);
        R a2 = N::max(N::max(DH, h_i[l]), h_north[l]);
        R u1 = divw<R>(qy1_north[l], a1);
        R u2 = divw<R>(qy1_i[l], a2);
        gy_north[l] = (u1 + u2 > 0) ? qy1_i[l] * u2 : qy1_north[l] * u1;
      }
      std::vector<R> fx_e(LANES), fy_w(LANES);
      for (int l = 0; l < LANES; l++) {
        fx_e[l] = down(fx_i, l);
        fy_w[l] = up(fy_i, l);
      }
      for (int l = 0; l < LANES; l++) {
        qx2_i[l] = rmg_qx<R>(fx_e[l], fx_i[l], gx_i[l], gx_south[l], dt, dx, dt_inverse,
                             dx_inverse, rb[l], h_i[l], h_east[l], sohle_east[l], sohle_i[l],
                             qx1_i[l], qy1_i[l], qy1_east[l], qy1_south[l], qy1_southeast[l]);
        qy2_i[l] = rmg_qy<R>(fy_i[l], fy_w[l], gy_i[l], gy_north[l], dt, dx, dt_inverse,
                             dx_inverse, rb[l], h_i[l], h_north[l], sohle_north[l], sohle_i[l],
                             qy1_i[l], qx1_i[l], qx1_west[l], qx1_north[l], qx1_northwest[l]);
        if (col[l] == 0 || col[l] == W - 2) qx2_i[l] = qx1_i[l];
        if (row == 0 || row == H - 2) qy2_i[l] = qy1_i[l];
      }
      for (int l = 0; l < LANES; l++) {
        if (!act[l]) continue;
        R qx2_west = up(qx2_i, l);
        R h2_i = h_i[l] - dt * dx_inverse * ((qx2_i[l] - qx2_west) + (qy2_i[l] - qy2_south[l]));
        const int c = col[l], tx = l;
        if (sohle_i[l] != s.no_data) {
          // quirks=1: the reference's store predicates verbatim (kernels.cu:386-405), which
          // also store warm-up rows 1 and H-2 (Q3, Q5) and halo lanes at col W-2 (Q4).
          // quirks=0: only rows whose south neighbours were computed (i >= 2, or row 1
          // after row 0) and only lanes whose east/west neighbours are in the warp.
          const bool valid_i = s.quirks ? true : (i >= 2 || (i == 1 && row0 + firstRow == 0));
          const bool innerRow = i >= STENCIL && row < H - STENCIL;
          const bool innerCol = tx >= STENCIL && tx < LANES - STENCIL && c < W - STENCIL;
          const bool r1 = (row == 1) && valid_i;
          const bool rH = (row == H - 2) && valid_i;
          const bool cW = (c == W - 2) && (s.quirks || (tx >= 2 && tx <= LANES - 2));
          if ((innerCol || c == 1 || cW) && (innerRow || r1 || rH)) {
            h2[gi + c] = h2_i;
            cells++;
            if (h2_i > EPS) {
              R umax = N::max(N::fabs(uy_i[l]),
                              N::max(N::fabs(uy_south[l]),
                                     N::max(N::fabs(ux_i[l]), N::fabs(ux_w[l]))));
              R dt_cfl = dx / (N::sqrt(G * h2_i) + umax);
              dtc[l] = N::min(dtc[l], dt_cfl);
              mnew[l] |= 0x8000000000000000ull;
            }
          }
          if ((innerCol || c == 1) && innerRow) qx2[gi + c] = qx2_i[l];
          if ((innerRow || r1) && innerCol) qy2[gi + c] = qy2_i[l];
        }
      }
      // travel north
      for (int l = 0; l < LANES; l++) {
        qy2_south[l] = qy2_i[l];
        qy1_south[l] = qy1_i[l];
        gx_south[l] = gx_i[l];
        qy1_southeast[l] = qy1_east[l];
        uy_south[l] = uy_i[l];
        h_i[l] = h_north[l];
        sohle_i[l] = sohle_north[l];
        rb[l] = rb_n[l];
        qx1_i[l] = qx1_north[l];
        qy1_i[l] = qy1_north[l];
        gy_i[l] = gy_north[l];
        h_east[l] = h_northeast[l];
        qx1_west[l] = qx1_northwest[l];
        h_north[l] = h_northnorth[l];
        qx1_north[l] = qx_nn[l];
        qy1_north[l] = qy_nn[l];
        sohle_north[l] = so_nn[l];
        h_northnorth[l] = h_nnn[l];
      }
      for (int l = 0; l < LANES; l++) {
        qy1_east[l] = down(qy1_i, l);
        sohle_east[l] = down(sohle_i, l);
        h_northeast[l] = down(h_north, l);
        qx1_northwest[l] = up(qx1_north, l);
      }
    }
    // store_courant (kernels.cu:173-192)
    {
      const int num_active_threads = wx * (32 - STENCIL * 2);
      const int boundary_distance = W - num_active_threads;
      R m = dtc[0];
      if (boundary_distance < 32) {
        for (int i = 1; i < boundary_distance; i++) m = N::min(m, act[i] ? dtc[i] : R(0));
      } else {
        for (int l = 1; l < LANES; l++) m = N::min(m, act[l] ? dtc[l] : R(0));
      }
      s.cfl_table[t] = m;
    }
    // store_mask (kernels.cu:156-171)
    if (s.warp_mask) {
      uint64_t mm = 0;
      for (int l = 0; l < LANES; l++) {
        if (!act[l] || totalrows <= 0) continue;
        uint64_t v = mnew[l];
        if (totalrows < 64) v >>= 64 - (totalrows + firstRow);
        mm |= v;
      }
      s.mask[t] = mm;
    }
  }
  s.cells_updated = cells;
  s.tiles_computed = tiles;
}

// ------------------------------------------------------------------------------------------
// Boundary conditions (src/bc_kernel_gen.cpp:11-152), applied in list order.
template <class R> void apply_one_bc(St<R>& s, const BC& bc, float qvalf, float zvalf) {
  typedef Num<R> N;
  const R G = K<R>::G;
  const R qval = (R)qvalf, zval = (R)zvalf;
  R* h = s.h[s.cur].data();
  R* qx = s.qx[s.cur].data();
  R* qy = s.qy[s.cur].data();
  const R* sohle = s.bed.data();
  const int W = s.W, H = s.H, pitch = s.P;
  const int ntid = (std::max(W, H) / 64 + 1) * 64;
  for (int tid = bc.from; tid <= bc.to && tid < bc.from + ntid; tid++) {
    long idx = 0, stag = 0, nb = 0;
    R* q = qx;
    R* ql = qy;
    if (bc.side == 1) {  // right
      idx = (long)tid * pitch + W - 1; stag = idx - 1; nb = idx - 1; q = qx; ql = qy;
    } else if (bc.side == 0) {  // left
      idx = (long)tid * pitch; stag = idx; nb = idx + 1; q = qx; ql = qy;
    } else if (bc.side == 2) {  // top
      idx = tid + (long)(H - 1) * pitch; stag = idx - pitch; nb = idx - pitch; q = qy; ql = qx;
    } else {  // bottom
      idx = tid; stag = idx; nb = idx + pitch; q = qy; ql = qx;
    }
    // open BC direction: +1 on the right/top side, -1 on the left/bottom side; the
    // reference's (float)(idx - idx_neighbor) is +-pitch on the top/bottom sides (Q8)
    R sign = (bc.side == 1 || bc.side == 2) ? R(1) : R(-1);
    if (s.open_pitch > 0 && bc.side >= 2) sign *= R(s.open_pitch);
    if (s.staggered) {
      const long qi = stag;
      if (bc.type == 0) {  // z
        h[idx] = N::max(R(0.f), zval - sohle[idx]);
        h[nb] = N::max(R(0.f), zval - sohle[nb]);
        q[qi] = R(-0.f);
      } else if (bc.type == 1) {  // q
        R ah = zval - sohle[idx];
        if (ah < R(0.f)) ah = R(0.f);
        q[qi] = N::sqrt(G * ah) * ah * qval;
        if (ah > R(0.f)) {
          R z_neighbor = h[nb] + sohle[nb];
          h[idx] = N::max(ah / R(2), z_neighbor - sohle[idx]);
        }
      } else if (bc.type == 2) {  // close
        q[qi] = R(0.f);
        ql[idx] = ql[nb];
        h[idx] = h[nb];
      } else {  // open
        R ah_neighbor = h[nb];
        R z_neighbor = ah_neighbor + sohle[nb];
        if (ah_neighbor > R(0.f)) {
          q[qi] = sign * N::sqrt(G * ah_neighbor) * ah_neighbor;
          h[idx] = N::min(ah_neighbor, N::max(R(0.f), z_neighbor - sohle[idx]));
        }
      }
    } else {
      if (bc.type == 0) {
        R ah = N::max(R(0.f), zval - sohle[idx]);
        h[idx] = ah;
        q[idx] = R(0.99f) * (q[nb] / N::max(h[nb], R(0.01f))) * ah;
      } else if (bc.type == 1) {
        R ah = zval - sohle[idx];
        if (ah < R(0.f)) ah = R(0.f);
        q[idx] = N::sqrt(G * ah) * ah * qval;
        if (ah > R(0.f)) h[idx] = N::max(ah / R(2), h[nb] + sohle[nb] - sohle[idx]);
      } else if (bc.type == 2) {
        q[idx] = R(0.f);
        ql[idx] = ql[nb];
        h[idx] = h[nb];
      } else {
        R ah_neighbor = h[nb];
        R z_neighbor = ah_neighbor + sohle[nb];
        if (ah_neighbor > R(0.f)) {
          q[idx] = sign * N::sqrt(G * ah_neighbor) * ah_neighbor;
          h[idx] = N::min(ah_neighbor, N::max(R(0.f), z_neighbor - sohle[idx]));
        }
      }
    }
  }
}

// BoundaryCondition::sumWaterDepths (src/boundary_condition.cpp:6-18): a sequential
// binary32 sum over boundary cells n in [from, to) of sqrt(G*ah)*ah, times dx.
template <class R> float sum_water_depths(const St<R>& s, const BC& bc, float z_soll) {
  auto nth = [&](int n) -> long {
    if (bc.side == 0) return (long)n * s.W;
    if (bc.side == 1) return (long)n * s.W + (s.W - 1);
    if (bc.side == 2) return n + (long)(s.H - 1) * s.W;
    return n;
  };
  float ah_sum = 0.f;
  const long inc = nth(1) - nth(0);
  for (long idx = nth(bc.from); idx < nth(bc.to); idx += inc) {
    float ah = z_soll - s.bed_host[idx];
    if (ah < 0.01f) ah = 0.f;
    ah_sum += std::sqrt(9.81f * ah) * ah;
  }
  ah_sum *= (float)s.dx;
  return ah_sum;
}

// getQ / getZ (src/boundary_condition.cpp:45-77) at the current duration.
template <class R> void bc_params(const St<R>& s, const BC& bc, float* qv, float* zv) {
  const float t = (float)s.duration;
  if (!bc.series) {
    *zv = bc.z;
    *qv = (bc.type != 1) ? bc.q : bc.q / sum_water_depths(s, bc, bc.z);
  } else {
    // a series BC without a z series reads z as 0 (the reference dereferences null)
    float z = bc.has_zs ? bc.zs.interp(t) : 0.f;
    *zv = z;
    if (bc.type != 1)
      *qv = 0.f;
    else
      *qv = (bc.has_qs ? bc.qs.interp(t) : 0.f) / sum_water_depths(s, bc, z);
  }
}

template <class R> void apply_bcs(St<R>& s) {
  for (const BC& bc : s.bcs) {
    float qv, zv;
    bc_params(s, bc, &qv, &zv);
    apply_one_bc(s, bc, qv, zv);
  }
}

// controlTimestep (src/main.cpp:25-43).
template <class R> void control(St<R>& s) {
  R min_cfl = R(42.0f);
  if (s.mode == ORC_MODE_WARP_REF) {
    for (R v : s.cfl_table)
      if (v < min_cfl) min_cfl = v;
  } else {
    if (s.cfl_acc < min_cfl) min_cfl = s.cfl_acc;
  }
  R target = min_cfl * R(.5f);
  const R dc = R(0.001f);
  if (s.dt < target && target - s.dt > dc)
    s.dt *= R(1.001f);
  else if (s.dt > target && s.dt - target > dc)
    s.dt *= R(0.99f);
}

template <class R> void stencil(St<R>& s) {
  if (s.mode == ORC_MODE_DENSE)
    stencil_dense(s);
  else if (s.mode == ORC_MODE_TILED)
    stencil_tiled(s);
  else
    stencil_warp_ref(s);
  s.duration += (double)s.dt;
  s.iteration++;
  s.cur ^= 1;
}

template <class R> void floodplain(St<R>& s) {
  const R* h = s.h[s.cur].data();
  for (int j = 0; j < s.H; j++)
    for (int i = 0; i < s.W; i++) {
      size_t g = s.idx(i, j);
      s.hmax[g] = Num<R>::max(s.hmax[g], h[g]);
    }
}

// simulate() (src/main.cpp:100-135) without sampling and dumps.
template <class R> void simulate_step(St<R>& s) {
  if (s.variable_dt) control(s);
  if (s.iteration == 0) apply_bcs(s);
  stencil(s);
  apply_bcs(s);
  if (s.iteration % 100 == 0) floodplain(s);
}

}  // namespace

struct orc_state {
  int prec;
  St<float>* f = nullptr;
  St<double>* d = nullptr;
};

template <class R>
static St<R>* make_state(const orc_desc* d, const float* bed, const float* h, const float* qx,
                         const float* qy, const float* kv) {
  St<R>* s = new St<R>();
  s->W = d->W;
  s->H = d->H;
  s->P = d->W;
  s->dx = (R)d->dx;
  s->dt = (R)d->dt;
  s->no_data = (R)d->no_data;
  s->kst = (R)d->kst;
  s->has_kvar = d->has_kst_var && kv;
  s->mode = d->mode;
  s->tw = d->tile_w;
  s->tr = d->tile_rows;
  s->warp_mask = d->warp_mask;
  s->quirks = d->quirks;
  s->open_pitch = d->open_pitch;
  s->variable_dt = d->variable_dt;
  s->staggered = d->staggered;
  s->threads = d->threads > 0 ? d->threads : 1;
#ifdef _OPENMP
  if (d->threads <= 0) s->threads = omp_get_max_threads();
#endif
  const size_t n = (size_t)s->P * (s->H + 3);
  auto fill = [&](std::vector<R>& v, const float* src) {
    v.assign(n, R(0));
    if (src)
      for (int j = 0; j < s->H; j++)
        for (int i = 0; i < s->W; i++) v[s->idx(i, j)] = (R)src[(size_t)j * s->W + i];
  };
  fill(s->bed, bed);
  s->bed_host.assign(bed, bed + (size_t)s->W * s->H);
  fill(s->kv, s->has_kvar ? kv : nullptr);
  for (int k = 0; k < 2; k++) {
    fill(s->h[k], h);
    fill(s->qx[k], qx);
    fill(s->qy[k], qy);
  }
  fill(s->hmax, h);
  if (s->mode == ORC_MODE_TILED) {
    s->tiles_x = (s->W + s->tw - 1) / s->tw;
    s->tiles_y = (s->H + s->tr - 1) / s->tr;
    s->mask.assign((size_t)s->tiles_x * s->tiles_y, 0);
    for (int ty = 0; ty < s->tiles_y; ty++)
      for (int tx = 0; tx < s->tiles_x; tx++)
        s->mask[(size_t)ty * s->tiles_x + tx] = valid_rows(std::min(s->tr, s->H - ty * s->tr));
  } else if (s->mode == ORC_MODE_WARP_REF) {
    const int cols = s->tw - 4;
    const int gx = s->W / (cols * 4) + 1, gy = s->H / s->tr + 1;
    s->tiles_x = gx * 4;
    s->tiles_y = gy;
    s->mask.assign((size_t)gx * 4 * gy, ~0ull);
    s->cfl_table.assign((size_t)gx * 4 * gy, R(1.f));
  }
  return s;
}

template <class R> static std::vector<R>* field_ptr(St<R>& S, int field, int side) {
  int k = side ? S.cur ^ 1 : S.cur;
  switch (field) {
    case 0: return &S.h[k];
    case 1: return &S.qx[k];
    case 2: return &S.qy[k];
    case 3: return &S.hmax;
    case 4: return &S.bed;
    default: return &S.kv;
  }
}

extern "C" {

orc_state* orc_create(const orc_desc* d, const float* bed, const float* h, const float* qx,
                      const float* qy, const float* kst_var) {
  orc_state* s = new orc_state();
  s->prec = d->precision;
  if (d->precision == 64)
    s->d = make_state<double>(d, bed, h, qx, qy, kst_var);
  else
    s->f = make_state<float>(d, bed, h, qx, qy, kst_var);
  return s;
}

void orc_destroy(orc_state* s) {
  delete s->f;
  delete s->d;
  delete s;
}

#define DISPATCH(expr) \
  do {                 \
    if (s->f) {        \
      auto& S = *s->f; \
      expr;            \
    } else {           \
      auto& S = *s->d; \
      expr;            \
    }                  \
  } while (0)

int orc_add_bc(orc_state* s, int type, int side, int from, int to, float z, float q,
               const float* zt, const float* zv, int zn, const float* qt, const float* qv,
               int qn) {
  BC bc;
  bc.type = type;
  bc.side = side;
  bc.from = from;
  bc.to = to;
  bc.z = z;
  bc.q = q;
  bc.has_zs = zn > 0;
  bc.has_qs = qn > 0;
  bc.series = bc.has_zs || bc.has_qs;
  if (bc.has_zs) {
    bc.zs.t.assign(zt, zt + zn);
    bc.zs.v.assign(zv, zv + zn);
  }
  if (bc.has_qs) {
    bc.qs.t.assign(qt, qt + qn);
    bc.qs.v.assign(qv, qv + qn);
  }
  DISPATCH(S.bcs.push_back(bc));
  return 0;
}

void orc_simulate_step(orc_state* s) { DISPATCH(simulate_step(S)); }
void orc_control_timestep(orc_state* s) { DISPATCH(control(S)); }
void orc_apply_bcs(orc_state* s) { DISPATCH(apply_bcs(S)); }
void orc_stencil(orc_state* s) { DISPATCH(stencil(S)); }
void orc_floodplain(orc_state* s) { DISPATCH(floodplain(S)); }

void orc_bc_params(orc_state* s, int i, float* qval, float* zval) {
  DISPATCH(bc_params(S, S.bcs[i], qval, zval));
}
float orc_sum_water_depths(orc_state* s, int i, float z) {
  float r = 0;
  DISPATCH(r = sum_water_depths(S, S.bcs[i], z));
  return r;
}
static long long bad_slot_counter_ = 0;

float orc_interp(const float* t, const float* v, int n, float time) {
  Series se;
  se.t.assign(t, t + n);
  se.v.assign(v, v + n);
  return se.interp(time);
}
float orc_pow233(float x) { return fip_pow233(x); }

// Every binary32 x with lo <= x <= hi against x^2.33f in long double (64-bit significand)
// rounded to binary32: the inputs counted and those where fip_pow233 differs, with up to
// `cap` of them written to bad_x.
void orc_pow233_sweep(float lo, float hi, long long* n, long long* nbad, float* bad_x, int cap) {
  const uint32_t a = fip_f2u(lo), b = fip_f2u(hi);
  long long cnt = 0, bad = 0;
  const long double p = (long double)2.33f;
  bad_slot_counter_ = 0;
#pragma omp parallel for schedule(static) reduction(+ : cnt, bad)
  for (long long u = a; u <= (long long)b; u++) {
    const float x = fip_u2f((uint32_t)u);
    const float want = (float)powl((long double)x, p);
    cnt++;
    if (fip_pow233(x) != want) {
      long long k;
#pragma omp atomic capture
      k = bad_slot_counter_++;
      if (k < cap) bad_x[k] = x;
      bad++;
    }
  }
  *n = cnt;
  *nbad = bad;
}

void orc_get_time(orc_state* s, double* t, double* dt, int* it) {
  DISPATCH({
    *t = S.duration;
    *dt = (double)S.dt;
    *it = S.iteration;
  });
}
void orc_set_dt(orc_state* s, double dt) { DISPATCH(S.dt = (decltype(S.dt))dt); }

void orc_get_field(orc_state* s, int field, int side, float* out) {
  DISPATCH({
    auto* v = field_ptr(S, field, side);
    for (int j = 0; j < S.H; j++)
      for (int i = 0; i < S.W; i++)
        out[(size_t)j * S.W + i] = v->empty() ? 0.f : (float)(*v)[S.idx(i, j)];
  });
}
void orc_get_field64(orc_state* s, int field, int side, double* out) {
  DISPATCH({
    auto* v = field_ptr(S, field, side);
    for (int j = 0; j < S.H; j++)
      for (int i = 0; i < S.W; i++)
        out[(size_t)j * S.W + i] = v->empty() ? 0.0 : (double)(*v)[S.idx(i, j)];
  });
}
void orc_set_field(orc_state* s, int field, int side, const float* in) {
  DISPATCH({
    auto* v = field_ptr(S, field, side);
    for (int j = 0; j < S.H; j++)
      for (int i = 0; i < S.W; i++) (*v)[S.idx(i, j)] = in[(size_t)j * S.W + i];
  });
}
int orc_num_tiles(orc_state* s) {
  int n = 0;
  DISPATCH(n = (int)S.mask.size());
  return n;
}
void orc_get_masks(orc_state* s, uint64_t* out) {
  DISPATCH(std::copy(S.mask.begin(), S.mask.end(), out));
}
double orc_cfl_min(orc_state* s) {
  double r = 0;
  DISPATCH({
    if (S.mode == ORC_MODE_WARP_REF) {
      double m = 42.0;
      for (auto v : S.cfl_table)
        if (v < m) m = v;
      r = m;
    } else {
      r = (double)S.cfl_acc;
    }
  });
  return r;
}
long long orc_cells_updated(orc_state* s) {
  long long r = 0;
  DISPATCH(r = S.cells_updated);
  return r;
}
long long orc_tiles_computed(orc_state* s) {
  long long r = 0;
  DISPATCH(r = S.tiles_computed);
  return r;
}

}  // extern "C"
