// Exact sheet substep for gfx950: one kernel per stage of the reference
// substep (src/fluid_core.rs:1879-2082, open and walled), each stage over one work
// box plus its 2-cell rim.  Arithmetic follows the reference operation by
// operation (no contraction: the file is built with -ffp-contract=off; IEEE
// division), so on the same inputs it reproduces the FP32 reference bit for bit
// wherever no libm transcendental enters (the damping factor is passed in when
// the viscosity is uniform).  This is the parity anchor of the library; the
// fused streaming kernel (sheet_fused.hip) is the throughput path.
//
// Scratch convention: inside the box every scratch value is the reference's;
// on the 2-cell rim the filtered planes are zero, as if each box started from
// a cleared scratch (the reference reuses a thread-local scratch whose stale
// rim values never reach a cell that is wet or next to a wet cell).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sheet_device.hpp"

namespace silt {

namespace {

constexpr float kGauss[5] = {0.06f, 0.24f, 0.40f, 0.24f, 0.06f};
constexpr float kSteep[5] = {0.02f, 0.08f, 0.80f, 0.08f, 0.02f};

struct Region {
  int i0, i1, j0, j1;
};

__device__ inline bool in_box(const Box& b, int i, int j) {
  return i >= b.i0 && i <= b.i1 && j >= b.j0 && j <= b.j1;
}

__device__ inline int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// accum5 / accum5_clamped (src/fluid_core.rs:3681-3698): interior taps start
// from the first product, clamped taps from 0.0.
__device__ inline float accum5(float a0, float a1, float a2, float a3, float a4, const float* k) {
  float a = a0 * k[0];
  a = a + a1 * k[1];
  a = a + a2 * k[2];
  a = a + a3 * k[3];
  a = a + a4 * k[4];
  return a;
}

__device__ inline float xblur_at(const float* src, int res, int i, int j, const float* k) {
  const float* row = src + (size_t)j * res;
  if (i >= 2 && i + 2 < res) return accum5(row[i - 2], row[i - 1], row[i], row[i + 1], row[i + 2], k);
  float a = 0.0f;
  for (int t = -2; t <= 2; t++) a = a + row[clampi(i + t, 0, res - 1)] * k[t + 2];
  return a;
}

__device__ inline float zblur_at(const float* src, int res, int i, int j, const float* k) {
  if (j >= 2 && j + 2 < res)
    return accum5(src[(size_t)(j - 2) * res + i], src[(size_t)(j - 1) * res + i], src[(size_t)j * res + i],
                  src[(size_t)(j + 1) * res + i], src[(size_t)(j + 2) * res + i], k);
  float a = 0.0f;
  for (int t = -2; t <= 2; t++) a = a + src[(size_t)clampi(j + t, 0, res - 1) * res + i] * k[t + 2];
  return a;
}

__device__ inline void grad_at(const float* p, int res, int i, int j, float cell, float* dx, float* dz) {
  int ip = i + 1 < res ? i + 1 : res - 1, im = i > 0 ? i - 1 : 0;
  int jp = j + 1 < res ? j + 1 : res - 1, jm = j > 0 ? j - 1 : 0;
  *dx = (p[(size_t)j * res + ip] - p[(size_t)j * res + im]) / (2.0f * cell);
  *dz = (p[(size_t)jp * res + i] - p[(size_t)jm * res + i]) / (2.0f * cell);
}

__device__ inline float vort_at(const float* u, const float* v, int res, int i, int j, float cell) {
  int ip = i + 1 < res ? i + 1 : res - 1, im = i > 0 ? i - 1 : 0;
  int jp = j + 1 < res ? j + 1 : res - 1, jm = j > 0 ? j - 1 : 0;
  float dvdx = (v[(size_t)j * res + ip] - v[(size_t)j * res + im]) / (2.0f * cell);
  float dudz = (u[(size_t)jp * res + i] - u[(size_t)jm * res + i]) / (2.0f * cell);
  return dvdx - dudz;
}

// Confinement kick from (su, sv) added to (u, v) at one cell
// (confine_one / confine_jacobi, src/fluid_core.rs:2402-2427, 3442-3471).
__device__ inline void confine_at(const float* su, const float* sv, float* u, float* v, int res, int i, int j,
                                  float cell, float dt, float eps) {
  float w = vort_at(su, sv, res, i, j, cell);
  if (!isfinite(w) || fabsf(w) < 1.0e-4f) return;
  float wxp = fabsf(vort_at(su, sv, res, i + 1 < res ? i + 1 : res - 1, j, cell));
  float wxm = fabsf(vort_at(su, sv, res, i > 0 ? i - 1 : 0, j, cell));
  float wzp = fabsf(vort_at(su, sv, res, i, j + 1 < res ? j + 1 : res - 1, cell));
  float wzm = fabsf(vort_at(su, sv, res, i, j > 0 ? j - 1 : 0, cell));
  float nx = (wxp - wxm) / (2.0f * cell);
  float nz = (wzp - wzm) / (2.0f * cell);
  float du = eps * dt * (-nz * w);
  float dv = eps * dt * (nx * w);
  if (isfinite(du) && isfinite(dv) && du * du + dv * dv <= 0.0625f) {
    size_t idx = (size_t)j * res + i;
    u[idx] = u[idx] + du;
    v[idx] = v[idx] + dv;
  }
}

// ---- stage kernels --------------------------------------------------------

// Launch helper: a 2D grid over an inclusive region.
__device__ inline bool region_cell(const Region& r, int* i, int* j) {
  *i = r.i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  *j = r.j0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  return *i <= r.i1 && *j <= r.j1;
}

__global__ void k_eta(const float* __restrict__ h, const float* __restrict__ bed, float* __restrict__ eta,
                      Region r, int res) {
  int i, j;
  if (!region_cell(r, &i, &j)) return;
  size_t idx = (size_t)j * res + i;
  eta[idx] = bed[idx] + h[idx];
}

// Both 5-tap filters along x; zero on the rim outside the box.
__global__ void k_xblur2(const float* __restrict__ src, float* __restrict__ tg, float* __restrict__ ts,
                         Region r, Box b, int res) {
  int i, j;
  if (!region_cell(r, &i, &j)) return;
  size_t idx = (size_t)j * res + i;
  if (in_box(b, i, j)) {
    tg[idx] = xblur_at(src, res, i, j, kGauss);
    ts[idx] = xblur_at(src, res, i, j, kSteep);
  } else {
    tg[idx] = 0.0f;
    ts[idx] = 0.0f;
  }
}

// Pressure d = 0.45 (0.62 G + 0.38 S) + 0.55 eta (src/fluid_core.rs:1923-1935).
__global__ void k_zblur_pressure(const float* __restrict__ tg, const float* __restrict__ ts,
                                 const float* __restrict__ eta, float* __restrict__ p, Region r, Box b, int res) {
  int i, j;
  if (!region_cell(r, &i, &j)) return;
  size_t idx = (size_t)j * res + i;
  if (in_box(b, i, j)) {
    float g = zblur_at(tg, res, i, j, kGauss);
    float s = zblur_at(ts, res, i, j, kSteep);
    float filtered = 0.62f * g + 0.38f * s;
    p[idx] = 0.45f * filtered + 0.55f * eta[idx];
  } else {
    p[idx] = 0.0f;
  }
}

// Correction potential q = 0.62 G(b) + 0.38 S(b) (src/fluid_core.rs:2021-2028).
__global__ void k_zblur_q(const float* __restrict__ tg, const float* __restrict__ ts, float* __restrict__ q,
                          Region r, Box b, int res) {
  int i, j;
  if (!region_cell(r, &i, &j)) return;
  size_t idx = (size_t)j * res + i;
  if (in_box(b, i, j)) {
    float g = zblur_at(tg, res, i, j, kGauss);
    float s = zblur_at(ts, res, i, j, kSteep);
    q[idx] = 0.62f * g + 0.38f * s;
  } else {
    q[idx] = 0.0f;
  }
}

// Gravity (or wind blend) on wet box cells, then the Jacobi snapshot on the
// box and its rim (force_gravity_par :2263-2319, gravity_jacobi_serial
// :2533-2570, force_wind_par :2684-2716).
__global__ void k_force(const float* __restrict__ h, float* __restrict__ u, float* __restrict__ v,
                        const float* __restrict__ nu, const float* __restrict__ p, float* __restrict__ su,
                        float* __restrict__ sv, Region r, Box b, StepArgs a) {
  int i, j;
  if (!region_cell(r, &i, &j)) return;
  int res = a.res;
  size_t idx = (size_t)j * res + i;
  float uu = u[idx], vv = v[idx];
  if (in_box(b, i, j) && h[idx] > a.eps) {
    if (a.wind) {
      uu = uu + (a.wx - uu) * a.blend;
      vv = vv + (a.wz - vv) * a.blend;
    } else {
      float px, pz;
      grad_at(p, res, i, j, a.cell, &px, &pz);
      float mob = a.nu_uniform ? a.mob0 : sheet_mobility(nu[idx]);
      uu = uu + -kGravity * px * mob * a.dt;
      vv = vv + -kGravity * pz * mob * a.dt;
    }
    u[idx] = uu;
    v[idx] = vv;
  }
  su[idx] = uu;
  sv[idx] = vv;
}

__global__ void k_confine_jacobi(const float* __restrict__ h, float* __restrict__ u, float* __restrict__ v,
                                 const float* __restrict__ su, const float* __restrict__ sv, Region r, StepArgs a) {
  int i, j;
  if (!region_cell(r, &i, &j)) return;
  size_t idx = (size_t)j * a.res + i;
  if (h[idx] <= a.eps) return;
  confine_at(su, sv, u, v, a.res, i, j, a.cell, a.dt, a.vort);
}

// Forward Gauss-Seidel sweep of gravity + confinement over one box
// (force_gravity_open :2236-2260, force_wind_open :2660-2681), scheduled as
// wavefronts t = (i - i0) + 2 (j - j0): every cell a kick reads that comes
// earlier in raster order has a smaller t, every later one a larger t, and
// no two cells of one wavefront read each other.  One workgroup.
__global__ void k_force_gauss_seidel(const float* __restrict__ h, float* u, float* v, const float* __restrict__ nu,
                                     const float* __restrict__ p, Box b, StepArgs a) {
  int res = a.res;
  int W = b.i1 - b.i0 + 1, H = b.j1 - b.j0 + 1;
  int tmax = (W - 1) + 2 * (H - 1);
  for (int t = 0; t <= tmax; t++) {
    // Cells with di + 2 dj = t: dj from max(0, ceil((t - (W-1))/2)) to min(H-1, t/2).
    int dj_lo = t - (W - 1) > 0 ? (t - (W - 1) + 1) / 2 : 0;
    int dj_hi = t / 2 < H - 1 ? t / 2 : H - 1;
    for (int dj = dj_lo + (int)threadIdx.x; dj <= dj_hi; dj += (int)blockDim.x) {
      int i = b.i0 + (t - 2 * dj), j = b.j0 + dj;
      size_t idx = (size_t)j * res + i;
      if (h[idx] <= a.eps) continue;
      if (a.wind) {
        u[idx] = u[idx] + (a.wx - u[idx]) * a.blend;
        v[idx] = v[idx] + (a.wz - v[idx]) * a.blend;
      } else {
        float px, pz;
        grad_at(p, res, i, j, a.cell, &px, &pz);
        float mob = a.nu_uniform ? a.mob0 : sheet_mobility(nu[idx]);
        u[idx] = u[idx] + -kGravity * px * mob * a.dt;
        v[idx] = v[idx] + -kGravity * pz * mob * a.dt;
      }
      if (a.vort > 0.0f) confine_at(u, v, u, v, res, i, j, a.cell, a.dt, a.vort);
    }
    __syncthreads();
  }
}

// Kicks of a box's wet cells, gravity or the wind blend (the first half of
// force_gravity_open :2236-2260 and force_wind_open :2660-2681, per cell),
// into (ku, kv).  A cell's kick reads p, its viscosity and its own u and v,
// and the in-place confinement sweep changes none of them before it reaches
// the cell (a cell's confinement writes only that cell), so every kick can
// come before the sweep, all cells at once.  ku, kv may be u, v: without
// confinement the kicks are the whole force pass.  Dry cells are left as
// they are (ku, kv unwritten).
__global__ void k_kick_box(const float* __restrict__ h, const float* u, const float* v, const float* __restrict__ nu,
                           const float* __restrict__ p, float* ku, float* kv, Region r, StepArgs a) {
  int i, j;
  if (!region_cell(r, &i, &j)) return;
  const int res = a.res;
  const size_t idx = (size_t)j * res + i;
  if (h[idx] <= a.eps) return;
  float uu = u[idx], vv = v[idx];
  if (a.wind) {
    uu = uu + (a.wx - uu) * a.blend;
    vv = vv + (a.wz - vv) * a.blend;
  } else {
    float px, pz;
    grad_at(p, res, i, j, a.cell, &px, &pz);
    const float mob = a.nu_uniform ? a.mob0 : sheet_mobility(nu[idx]);
    uu = uu + -kGravity * px * mob * a.dt;
    vv = vv + -kGravity * pz * mob * a.dt;
  }
  ku[idx] = uu;
  kv[idx] = vv;
}

// The confinement half of the forward Gauss-Seidel sweep, with the kicks
// already in (ku, kv) (k_kick_box) and the planes it reads held in LDS.  At
// wavefront t ring row jj updates column c(jj, t) = i0 + t - 2 (jj - j0);
// every cell of that wavefront reads row jj only within c +- 4 (the
// vorticity stencils reach 2 columns in their own row, 1 column one row
// away, the same column two rows away).  One lane per ring row (box rows
// plus 2 above and below, clamped to the grid); each row keeps a ring of 32
// columns of u and v, and the cell's own depth and kicked velocity for the
// chunk in LDS.
//
// Memory moves kChunk wavefronts at a time and coalesced: a lane's row is
// far from its neighbours' in memory, so a load or store per lane and
// wavefront costs the CU's memory pipeline one cache line per lane; instead
// each copy instruction covers 8 rows x 8 consecutive columns of the
// wave's rows.  Between two chunks, after a barrier: the columns loaded
// during the chunk (c + 4 .. c + kChunk + 3 of the next chunk's first
// wavefront, not yet updated: a row's cells change only when the wavefront
// reaches them) go to the ring, the chunk before the last goes back to
// memory, and the next chunk's columns are loaded into registers.  The ring
// slots a chunk's columns land in held columns at least 13 behind the
// wavefront, which no cell reads any more.  Same arithmetic in the same
// order as the global-memory sweep: the confinement of a cell reads its own
// kicked values and its neighbours' values as the raster order leaves them
// (final before it, unkicked after it).  Needs the box's rows + 4 <= 256
// and vort > 0.
constexpr int kRing = 32;
constexpr int kRingPad = kRing + 1;  // lane = row: an odd row stride spreads the lanes over the banks
constexpr int kRingRows = 256;
constexpr int kChunk = 8;
static_assert(3 * kChunk + 3 < kRing, "the landed, current and flushed columns must fit the ring");

__global__ __launch_bounds__(256) void k_confine_gs_lds(const float* __restrict__ h, float* u, float* v,
                                                        const float* __restrict__ ku, const float* __restrict__ kv,
                                                        Box b, StepArgs a) {
  __shared__ float ru[kRingRows][kRingPad], rv[kRingRows][kRingPad];
  __shared__ float oh_[kRingRows][kChunk + 1], ou_[kRingRows][kChunk + 1], ov_[kRingRows][kChunk + 1];
  const int res = a.res;
  const int r_lo = b.j0 - 2 > 0 ? b.j0 - 2 : 0;
  const int r_hi = b.j1 + 2 < res - 1 ? b.j1 + 2 : res - 1;
  const int row = (int)threadIdx.x;
  const int jj = r_lo + row;
  const bool live = jj <= r_hi;
  const int W = b.i1 - b.i0 + 1, H = b.j1 - b.j0 + 1;
  const int tmax = (W - 1) + 2 * (H - 1);
  // x / (2 cell) is div_by(x, y2c): the IEEE quotient in three instructions.
  const double y2c = 1.0 / (double)(2.0f * a.cell);
  auto fast = [&](float x) -> float { return div_by(x, y2c); };
  const float kc = a.vort * a.dt;
  auto col_at = [&](int r, int t) { return b.i0 + t - 2 * (r - b.j0); };
  auto in_row = [&](int r, int c) { return r_lo + r <= r_hi && c >= 0 && c < res; };
  auto in_box = [&](int r, int c) {
    const int y = r_lo + r;
    return y <= r_hi && y >= b.j0 && y <= b.j1 && c >= b.i0 && c <= b.i1;
  };
  // Copy m of a lane covers ring row wbase + 8 m + lr, column l8 of the 8.
  const int wbase = (int)(threadIdx.x & ~63u), lr = (int)((threadIdx.x & 63u) >> 3), l8 = (int)(threadIdx.x & 7u);
  float qu[kChunk], qv[kChunk], qh[kChunk], qku[kChunk], qkv[kChunk];
  // The chunk after the one starting at wavefront t0: ring columns and own
  // values into registers.
  auto fetch = [&](int t0) {
#pragma unroll
    for (int m = 0; m < kChunk; m++) {
      const int r = wbase + 8 * m + lr, y = r_lo + r, cb = col_at(y, t0);
      const int c = cb + kChunk + 4 + l8;
      if (in_row(r, c)) {
        const size_t idx = (size_t)y * res + c;
        qu[m] = u[idx];
        qv[m] = v[idx];
      }
      const int co = cb + kChunk + l8;
      qh[m] = 0.0f;
      if (in_box(r, co)) {
        const size_t idx = (size_t)y * res + co;
        qh[m] = h[idx];
        qku[m] = ku[idx];
        qkv[m] = kv[idx];
      }
    }
  };
  // The registers to the ring and the own-value rows for the chunk starting
  // at wavefront t0 (own depth 0 outside the box: eps >= 0 skips the cell).
  auto land = [&](int t0) {
#pragma unroll
    for (int m = 0; m < kChunk; m++) {
      const int r = wbase + 8 * m + lr, cb = col_at(r_lo + r, t0);
      const int c = cb + 4 + l8;
      if (in_row(r, c)) {
        ru[r][c & (kRing - 1)] = qu[m];
        rv[r][c & (kRing - 1)] = qv[m];
      }
      oh_[r][l8] = qh[m];
      ou_[r][l8] = qku[m];
      ov_[r][l8] = qkv[m];
    }
  };
  // The final values of the chunk starting at wavefront t0 to memory.
  auto flush = [&](int t0) {
#pragma unroll
    for (int m = 0; m < kChunk; m++) {
      const int r = wbase + 8 * m + lr, y = r_lo + r, c = col_at(y, t0) + l8;
      if (in_box(r, c)) {
        const size_t idx = (size_t)y * res + c;
        u[idx] = ru[r][c & (kRing - 1)];
        v[idx] = rv[r][c & (kRing - 1)];
      }
    }
  };
  // Ring slot of grid cell (x, y), the reads clamped to the grid as the
  // global-memory helpers clamp them.
  auto at = [&](float (*rg)[kRingPad], int x, int y) -> float {
    x = x < 0 ? 0 : (x > res - 1 ? res - 1 : x);
    y = y < 0 ? 0 : (y > res - 1 ? res - 1 : y);
    return rg[y - r_lo][x & (kRing - 1)];
  };
  auto vort = [&](int x, int y) -> float {
    const int xp = x + 1 < res ? x + 1 : res - 1, xm = x > 0 ? x - 1 : 0;
    const int yp = y + 1 < res ? y + 1 : res - 1, ym = y > 0 ? y - 1 : 0;
    return fast(at(rv, xp, y) - at(rv, xm, y)) - fast(at(ru, x, yp) - at(ru, x, ym));
  };
  {
    // The first chunk, row by row.
    const int c0 = col_at(jj, 0);
    for (int c = c0 - 4; c <= c0 + kChunk + 3; c++)
      if (in_row(row, c)) {
        const size_t idx = (size_t)jj * res + c;
        ru[row][c & (kRing - 1)] = u[idx];
        rv[row][c & (kRing - 1)] = v[idx];
      }
    for (int k = 0; k < kChunk; k++) {
      const int c = c0 + k;
      oh_[row][k] = 0.0f;
      if (in_box(row, c)) {
        const size_t idx = (size_t)jj * res + c;
        oh_[row][k] = h[idx];
        ou_[row][k] = ku[idx];
        ov_[row][k] = kv[idx];
      }
    }
  }
  if (kChunk <= tmax) fetch(0);
  for (int t0 = 0;; t0 += kChunk) {
    const int t1 = t0 + kChunk - 1 < tmax ? t0 + kChunk - 1 : tmax;
    float nh = 0.0f, nu_ = 0.0f, nv_ = 0.0f;
    for (int t = t0; t <= t1; t++) {
      const int c = live ? col_at(jj, t) : 0;
      __syncthreads();
      // The cell's own values: the chunk's first from LDS now, the others
      // read one wavefront early.
      float ch = nh, cu = nu_, cv = nv_;
      if (t == t0) {
        ch = oh_[row][0];
        cu = ou_[row][0];
        cv = ov_[row][0];
      }
      if (t < t1) {
        nh = oh_[row][t + 1 - t0];
        nu_ = ou_[row][t + 1 - t0];
        nv_ = ov_[row][t + 1 - t0];
      }
      if (!(ch > a.eps)) continue;
      const int i = c, j = jj;
      float uu = cu, vv = cv;
      // Two cells or more from every grid side, no read clamps: every value
      // the cell needs is read from the ring first (none of them changes
      // during this wavefront), the cell's own values stay in registers, and
      // the confinement's tests become selects.
      if (i >= 2 && i <= res - 3 && j >= 2 && j <= res - 3) {
        const int y0 = j - r_lo;
        auto U = [&](int dx, int dy) { return ru[y0 + dy][(i + dx) & (kRing - 1)]; };
        auto Vv = [&](int dx, int dy) { return rv[y0 + dy][(i + dx) & (kRing - 1)]; };
        const float v_e = Vv(1, 0), v_w = Vv(-1, 0), v_ee = Vv(2, 0), v_ww = Vv(-2, 0);
        const float v_ne = Vv(1, 1), v_nw = Vv(-1, 1), v_se = Vv(1, -1), v_sw = Vv(-1, -1);
        const float u_n = U(0, 1), u_s = U(0, -1), u_nn = U(0, 2), u_ss = U(0, -2);
        const float u_ne = U(1, 1), u_se = U(1, -1), u_nw = U(-1, 1), u_sw = U(-1, -1);
        const float w = fast(v_e - v_w) - fast(u_n - u_s);
        const float wxp = fabsf(fast(v_ee - vv) - fast(u_ne - u_se));
        const float wxm = fabsf(fast(vv - v_ww) - fast(u_nw - u_sw));
        const float wzp = fabsf(fast(v_ne - v_nw) - fast(u_nn - uu));
        const float wzm = fabsf(fast(v_se - v_sw) - fast(uu - u_ss));
        const float nx = fast(wxp - wxm);
        const float nz = fast(wzp - wzm);
        const float du = kc * (-nz * w);
        const float dv = kc * (nx * w);
        const bool take = isfinite(w) && !(fabsf(w) < 1.0e-4f) && isfinite(du) && isfinite(dv) &&
                          du * du + dv * dv <= 0.0625f;
        uu = take ? uu + du : uu;
        vv = take ? vv + dv : vv;
      } else {
        // Near a grid side a clamped read may land on the cell itself: its
        // kicked values go to the ring first (no other cell of this
        // wavefront reads them).
        ru[row][i & (kRing - 1)] = uu;
        rv[row][i & (kRing - 1)] = vv;
        const float w = vort(i, j);
        if (isfinite(w) && !(fabsf(w) < 1.0e-4f)) {
          const float wxp = fabsf(vort(i + 1 < res ? i + 1 : res - 1, j));
          const float wxm = fabsf(vort(i > 0 ? i - 1 : 0, j));
          const float wzp = fabsf(vort(i, j + 1 < res ? j + 1 : res - 1));
          const float wzm = fabsf(vort(i, j > 0 ? j - 1 : 0));
          const float nx = fast(wxp - wxm);
          const float nz = fast(wzp - wzm);
          const float du = kc * (-nz * w);
          const float dv = kc * (nx * w);
          if (isfinite(du) && isfinite(dv) && du * du + dv * dv <= 0.0625f) {
            uu = uu + du;
            vv = vv + dv;
          }
        }
      }
      ru[row][i & (kRing - 1)] = uu;
      rv[row][i & (kRing - 1)] = vv;
    }
    // Every cell of the chunk is in the ring before any lane copies rows
    // that are not its own.
    __syncthreads();
    if (t0 + kChunk > tmax) {
      if (t0 >= kChunk) flush(t0 - kChunk);
      flush(t0);
      break;
    }
    land(t0 + kChunk);
    if (t0 >= kChunk) flush(t0 - kChunk);
    if (t0 + 2 * kChunk <= tmax) fetch(t0 + kChunk);
  }
}

// Divergence source b = div(u,v) ov - 0.015 ov lap(eta) (:1999-2000).
__global__ void k_project_source(const float* __restrict__ u, const float* __restrict__ v,
                                 const float* __restrict__ eta, const float* __restrict__ nu, float* __restrict__ bsrc,
                                 Region r, Box b, StepArgs a) {
  int i, j;
  if (!region_cell(r, &i, &j)) return;
  int res = a.res;
  size_t idx = (size_t)j * res + i;
  if (!in_box(b, i, j)) {
    bsrc[idx] = 0.0f;
    return;
  }
  float cell = a.cell;
  int ip = i + 1 < res ? i + 1 : res - 1, im = i > 0 ? i - 1 : 0;
  int jp = j + 1 < res ? j + 1 : res - 1, jm = j > 0 ? j - 1 : 0;
  float ux = (u[(size_t)j * res + ip] - u[(size_t)j * res + im]) / (2.0f * cell);
  float vz = (v[(size_t)jp * res + i] - v[(size_t)jm * res + i]) / (2.0f * cell);
  float div = ux + vz;
  float s = eta[(size_t)j * res + ip] + eta[(size_t)j * res + im] + eta[(size_t)jp * res + i] +
            eta[(size_t)jm * res + i] - 4.0f * eta[idx];
  float lap_h = s / (cell * cell);
  float overload = a.nu_uniform ? a.ov0 : overload_of(nu[idx]);
  bsrc[idx] = div * overload - 0.015f * overload * lap_h;
}

// Correction kick, damping and the speed cap on wet box cells (:2010-2042).
__global__ void k_damp(const float* __restrict__ h, float* __restrict__ u, float* __restrict__ v,
                       const float* __restrict__ nu, const float* __restrict__ q, Region r, StepArgs a) {
  int i, j;
  if (!region_cell(r, &i, &j)) return;
  size_t idx = (size_t)j * a.res + i;
  if (h[idx] <= a.eps) return;
  float uu = u[idx], vv = v[idx];
  if (a.project) {
    float qx, qz;
    grad_at(q, a.res, i, j, a.cell, &qx, &qz);
    uu = uu + 0.012f * a.dt * qx;
    vv = vv + 0.012f * a.dt * qz;
  }
  float decay = a.nu_uniform ? a.decay0 : expf(-sheet_damp(nu[idx]) * a.dt);
  u[idx] = finite_speed(uu * decay, a.lim);
  v[idx] = finite_speed(vv * decay, a.lim);
}

// open_edge_dh (src/fluid_core.rs:2887-2911).
__device__ inline float edge_dh(float hd, float hr, float remd, uint8_t gate_r, bool banks, float bed_d,
                                float bed_r, float speed, float dt, float cell) {
  bool live = !(fabsf(speed) < 1.0e-5f);
  bool bank = banks && hr <= 1.0e-6f && bed_r > bed_d + hd + 1.0e-3f;
  float room = hd - remd;
  bool gate_shut = gate_r == 0 && hr <= 1.0e-6f;
  float frac = fminf(fabsf(speed) * dt / cell, 0.45f);
  float dh = fminf(hd * frac, room);
  bool ok = live && !bank && !(room <= 1.0e-8f) && !gate_shut && !(dh <= 1.0e-8f);
  return ok ? dh : 0.0f;
}

// One open transport pass for one cell: its two box-interior edges along the
// axis, in the reference's summation order (transport_open_rows_body
// :2998-3137).  stride is 1 along x, res along z; lo/hi bound the cell's index
// along the axis inside the box.
__device__ inline float transport_cell(const float* hin, const float* vel, const uint8_t* gate,
                                       const float* bed, bool banks, size_t idx, size_t stride, int pos, int lo,
                                       int hi, float dt, float cell, bool gate_uniform) {
  auto gate_of = [&](size_t k) -> uint8_t { return gate_uniform ? (uint8_t)1 : gate[k]; };
  auto bed_of = [&](size_t k) -> float { return banks ? bed[k] : 0.0f; };
  // Left-going flux of edge (k, k+1): donor k+1.
  auto fl = [&](size_t k) -> float {
    float speed = 0.5f * (vel[k] + vel[k + stride]);
    float dh = edge_dh(hin[k + stride], hin[k], 0.0f, gate_of(k), banks, bed_of(k + stride), bed_of(k), speed, dt,
                       cell);
    return speed > 0.0f ? 0.0f : dh;
  };
  // Right-going flux of edge (k, k+1): donor k, whose removed sum holds the
  // left-going flux of the edge before it.
  auto fr = [&](size_t k, bool has_prev) -> float {
    float speed = 0.5f * (vel[k] + vel[k + stride]);
    float prev = has_prev ? fl(k - stride) : 0.0f;
    float dh = edge_dh(hin[k], hin[k + stride], 0.0f + prev, gate_of(k + stride), banks, bed_of(k),
                       bed_of(k + stride), speed, dt, cell);
    return speed > 0.0f ? dh : 0.0f;
  };
  bool has_left = pos > lo, has_right = pos < hi;
  float fl_prev = has_left ? fl(idx - stride) : 0.0f;
  float fr_prev = has_left ? fr(idx - stride, pos - 1 > lo) : 0.0f;
  float fl_here = has_right ? fl(idx) : 0.0f;
  float fr_here = has_right ? fr(idx, has_left) : 0.0f;
  float rem = (0.0f + fl_prev) + fr_here;
  float add = (0.0f + fr_prev) + fl_here;
  float hv = hin[idx];
  float moved = fmaxf(hv - rem + add, 0.0f);
  return (rem > 0.0f || add > 0.0f) ? moved : hv;
}

// Adds v summed over the 64 lanes of each wave to *p, one atomic per wave
// (the blocks here are 64 x 4: a wave is one row of the block).  Every lane
// of the block must reach it.
__device__ inline void book_wave(double* p, double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (threadIdx.x == 0 && v != 0.0) atomicAdd(p, v);
}

// One transport pass over the box.  The z pass (along_x 0) writes h over the
// depth the x pass started from, so with a ledger it books each cell's change
// across both passes: the exchanged face fluxes cancel over the grid and the
// sum is what FP32 rounding and the max(0) clamp created or removed.
__global__ void k_transport(const float* __restrict__ hin, float* __restrict__ hout, const float* __restrict__ vel,
                            const uint8_t* __restrict__ gate, const float* __restrict__ bed, Box b, int along_x,
                            StepArgs a, float dt) {
  Region r{b.i0, b.i1, b.j0, b.j1};
  int i, j;
  const bool in = region_cell(r, &i, &j);
  double moved = 0.0;
  if (in) {
    size_t idx = (size_t)j * a.res + i;
    if (along_x) {
      hout[idx] = transport_cell(hin, vel, gate, bed, a.banks, idx, 1, i, b.i0, b.i1, dt, a.cell, a.gate_uniform);
    } else {
      const float was = hout[idx];
      const float now = transport_cell(hin, vel, gate, bed, a.banks, idx, (size_t)a.res, j, b.j0, b.j1, dt, a.cell,
                                       a.gate_uniform);
      hout[idx] = now;
      moved = (double)fmaxf(now, 0.0f) - (double)fmaxf(was, 0.0f);
    }
  }
  if (!along_x && a.book) book_wave(&a.book->residue, moved);
}

// Dry cut (:2072-2081): depth under eps, and its velocity, go to zero.  With
// a ledger the deleted depth is booked as cut.
__global__ void k_dry_cut(float* __restrict__ h, float* __restrict__ u, float* __restrict__ v, Box b, StepArgs a) {
  Region r{b.i0, b.i1, b.j0, b.j1};
  int i, j;
  double cut = 0.0;
  if (region_cell(r, &i, &j)) {
    size_t idx = (size_t)j * a.res + i;
    const float hv = h[idx];
    if (a.eps > 0.0f && hv < a.eps) {
      cut = (double)fmaxf(hv, 0.0f);
      h[idx] = 0.0f;
      u[idx] = 0.0f;
      v[idx] = 0.0f;
    }
  }
  if (a.book) book_wave(&a.book->cut, cut);
}

// ---- walled sheet (hydro solids) -------------------------------------------
// The stages the substep adds beside walls (src/fluid_core.rs:1879-2082 with
// solid = Some): the surface lift and the no-slip ghost before the forces,
// the kick's mobility bump beside a wall, the redirect, the no-penetration
// filter and the ghost again after the damping, and the transport that no
// face between a wall and the sheet crosses.  Every stage keeps the
// reference's operation order; the collect-then-write stages run in place
// because each reads only cells it does not write.

__device__ inline bool grid_cell(int i, int j, int res) { return i >= 0 && j >= 0 && i < res && j < res; }

constexpr int kNb4[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};

// lift_solid_surface (:4225-4276): a solid box cell's eta becomes the mean
// eta of the wet open cells within two cells.  It reads open cells only and
// writes solid cells only.
__global__ void k_lift_solid(float* __restrict__ eta, const float* __restrict__ h, const uint8_t* __restrict__ solid,
                             Region r, int res, float eps) {
  int i, j;
  if (!region_cell(r, &i, &j)) return;
  const size_t idx = (size_t)j * res + i;
  if (!solid[idx]) return;
  float acc = 0.0f, n = 0.0f;
  for (int dj = -2; dj <= 2; dj++)
    for (int di = -2; di <= 2; di++) {
      if (di == 0 && dj == 0) continue;
      const int ni = i + di, nj = j + dj;
      if (!grid_cell(ni, nj, res)) continue;
      const size_t nidx = (size_t)nj * res + ni;
      if (solid[nidx] || h[nidx] <= eps) continue;
      acc = acc + eta[nidx];
      n = n + 1.0f;
    }
  if (n > 0.0f) eta[idx] = acc / n;
}

// write_noslip_ghost (:4278-4326): a solid box cell's velocity becomes minus
// the mean of its wet open 4-neighbours' (0 without one).
__global__ void k_noslip_ghost(float* __restrict__ u, float* __restrict__ v, const float* __restrict__ h,
                               const uint8_t* __restrict__ solid, Region r, int res, float eps) {
  int i, j;
  if (!region_cell(r, &i, &j)) return;
  const size_t idx = (size_t)j * res + i;
  if (!solid[idx]) return;
  float au = 0.0f, av = 0.0f, n = 0.0f;
  for (int k = 0; k < 4; k++) {
    const int ni = i + kNb4[k][0], nj = j + kNb4[k][1];
    if (!grid_cell(ni, nj, res)) continue;
    const size_t nidx = (size_t)nj * res + ni;
    if (solid[nidx] || h[nidx] <= eps) continue;
    au = au + u[nidx];
    av = av + v[nidx];
    n = n + 1.0f;
  }
  u[idx] = n > 0.0f ? -au / n : 0.0f;
  v[idx] = n > 0.0f ? -av / n : 0.0f;
}

// solid_near with radius 1 (:4565-4582).
__device__ inline bool solid_near1(const uint8_t* solid, int res, int i, int j) {
  for (int dj = -1; dj <= 1; dj++)
    for (int di = -1; di <= 1; di++) {
      if (di == 0 && dj == 0) continue;
      const int ni = i + di, nj = j + dj;
      if (grid_cell(ni, nj, res) && solid[(size_t)nj * res + ni]) return true;
    }
  return false;
}

// The gravity kick beside walls (:2118-2124): mobility at least 1 on a deep
// cell next to a wall.
__device__ inline void kick_walled(const float* h, float* u, float* v, const float* nu, const float* p,
                                   const uint8_t* solid, int res, int i, int j, const StepArgs& a) {
  const size_t idx = (size_t)j * res + i;
  float px, pz;
  grad_at(p, res, i, j, a.cell, &px, &pz);
  float mob = a.nu_uniform ? a.mob0 : sheet_mobility(nu[idx]);
  if (h[idx] > 0.08f && solid_near1(solid, res, i, j)) mob = fmaxf(mob, 1.0f);
  u[idx] = u[idx] + -kGravity * px * mob * a.dt;
  v[idx] = v[idx] + -kGravity * pz * mob * a.dt;
}

// force_gravity_solid_par's kicks (:2156-2185), every wet box cell at once.
__global__ void k_kick_walled(const float* __restrict__ h, float* __restrict__ u, float* __restrict__ v,
                              const float* __restrict__ nu, const float* __restrict__ p,
                              const uint8_t* __restrict__ solid, Region r, StepArgs a) {
  int i, j;
  if (!region_cell(r, &i, &j)) return;
  if (h[(size_t)j * a.res + i] <= a.eps) return;
  kick_walled(h, u, v, nu, p, solid, a.res, i, j, a);
}

// The forward sweep beside walls, one workgroup, wavefronts t = di + 2 dj
// (the schedule of k_force_gauss_seidel): kick (with the wall bump) and
// confinement cell by cell (force_gravity_solid :2101-2128), or, with kick
// 0, the confinement pass that follows the parallel kicks (:2187-2199).
__global__ void k_sweep_walled(const float* __restrict__ h, float* u, float* v, const float* __restrict__ nu,
                               const float* __restrict__ p, const uint8_t* __restrict__ solid, Box b, StepArgs a,
                               int kick) {
  const int res = a.res;
  const int W = b.i1 - b.i0 + 1, H = b.j1 - b.j0 + 1;
  const int tmax = (W - 1) + 2 * (H - 1);
  for (int t = 0; t <= tmax; t++) {
    const int dj_lo = t - (W - 1) > 0 ? (t - (W - 1) + 1) / 2 : 0;
    const int dj_hi = t / 2 < H - 1 ? t / 2 : H - 1;
    for (int dj = dj_lo + (int)threadIdx.x; dj <= dj_hi; dj += (int)blockDim.x) {
      const int i = b.i0 + (t - 2 * dj), j = b.j0 + dj;
      if (h[(size_t)j * res + i] <= a.eps) continue;
      if (kick) kick_walled(h, u, v, nu, p, solid, res, i, j, a);
      if (a.vort > 0.0f) confine_at(u, v, u, v, res, i, j, a.cell, a.dt, a.vort);
    }
    __syncthreads();
  }
}

// redirect_into_walls (:4446-4563), first half: each wet open box cell with
// h >= 0.02 walks its four walls in order and books, per wall, the share
// each open flank gets (share[k], k the wall's neighbour index) and which
// flanks take it (bits 2k and 2k+1 of flanks, as a small float), and its own
// removed depth.
__global__ void k_redirect_give(const float* __restrict__ h, const float* __restrict__ u,
                                const float* __restrict__ v, const uint8_t* __restrict__ solid,
                                const float* __restrict__ bed, Region r, StepArgs a, float* __restrict__ removed,
                                float* __restrict__ s0, float* __restrict__ s1, float* __restrict__ s2,
                                float* __restrict__ s3, float* __restrict__ flanks) {
  int i, j;
  if (!region_cell(r, &i, &j)) return;
  const int res = a.res;
  const size_t idx = (size_t)j * res + i;
  float share[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  int bits = 0;
  float rem = 0.0f;
  const float hv = h[idx];
  if (!solid[idx] && !(hv < 0.02f)) {
    for (int k = 0; k < 4; k++) {
      const int di = kNb4[k][0], dj = kNb4[k][1];
      const int ni = i + di, nj = j + dj;
      if (!grid_cell(ni, nj, res) || !solid[(size_t)nj * res + ni]) continue;
      float vn = u[idx] * (float)di + v[idx] * (float)dj;
      const int ui = i - di, uj = j - dj;
      if (grid_cell(ui, uj, res)) {
        const size_t uidx = (size_t)uj * res + ui;
        if (!solid[uidx]) {
          const float feed = u[uidx] * (float)di + v[uidx] * (float)dj;
          if (feed > vn) vn = feed;
        }
      }
      if (vn < 0.05f) continue;
      int nd = 0, kb = 0;
      float flank = 0.0f;
      for (int q = 0; q < 2; q++) {
        const int pi = di != 0 ? 0 : (q == 0 ? -1 : 1), pj = di != 0 ? (q == 0 ? -1 : 1) : 0;
        const int qi = i + pi, qj = j + pj;
        if (!grid_cell(qi, qj, res)) continue;
        const size_t qq = (size_t)qj * res + qi;
        if (solid[qq]) continue;
        if (h[qq] <= 1.0e-6f && bed[qq] > bed[idx] + hv + 1.0e-3f) continue;
        kb |= 1 << q;
        flank = flank + h[qq];
        nd++;
      }
      if (nd == 0) continue;
      const float excess = hv - flank / (float)nd;
      if (excess < 2.0e-3f) continue;
      const float room = hv - rem;
      if (room < 1.0e-6f) continue;
      const float frac = fminf(vn * a.dt / a.cell, 0.45f);
      const float dh = fminf(fminf(excess, hv * frac), room);
      if (dh < 1.0e-8f) continue;
      rem = rem + dh;
      share[k] = dh / (float)nd;
      bits |= kb << (2 * k);
    }
  }
  removed[idx] = rem;
  s0[idx] = share[0];
  s1[idx] = share[1];
  s2[idx] = share[2];
  s3[idx] = share[3];
  flanks[idx] = (float)bits;
}

// Second half, over the box and a 1-cell rim: each cell adds what its
// neighbours sent it in the reference's order (donors in raster order, each
// donor's walls in neighbour order), then h = max(h - removed + added, 0).
// With a ledger each cell's change is booked as residue.
__global__ void k_redirect_take(float* __restrict__ h, Region rim, Box b, StepArgs a,
                                const float* __restrict__ removed, const float* __restrict__ s0,
                                const float* __restrict__ s1, const float* __restrict__ s2,
                                const float* __restrict__ s3, const float* __restrict__ flanks) {
  int i, j;
  const bool in = region_cell(rim, &i, &j);
  double moved = 0.0;
  if (in) {
    const int res = a.res;
    const size_t idx = (size_t)j * res + i;
    const float* sh[4] = {s0, s1, s2, s3};
    float add = 0.0f;
    // Donor at (i + ox, j + oz): the walls k0, k1 whose flank bit q is this cell.
    auto from = [&](int ox, int oz, int k0, int k1, int q) {
      const int di = i + ox, dj = j + oz;
      if (!in_box(b, di, dj)) return;
      const size_t d = (size_t)dj * res + di;
      const int bits = (int)flanks[d];
      if (bits & (1 << (2 * k0 + q))) add = add + sh[k0][d];
      if (bits & (1 << (2 * k1 + q))) add = add + sh[k1][d];
    };
    from(0, -1, 0, 1, 1);  // the donor above: this cell is the +z flank of its x walls
    from(-1, 0, 2, 3, 1);  // the donor to the left: +x flank of its z walls
    from(1, 0, 2, 3, 0);   // the donor to the right: -x flank of its z walls
    from(0, 1, 0, 1, 0);   // the donor below: -z flank of its x walls
    const float rem = in_box(b, i, j) ? removed[idx] : 0.0f;
    if (rem > 0.0f || add > 0.0f) {
      const float was = h[idx];
      const float now = fmaxf(was - rem + add, 0.0f);
      h[idx] = now;
      moved = (double)now - (double)fmaxf(was, 0.0f);
    }
  }
  if (a.book) book_wave(&a.book->residue, moved);
}

// block_into_solid (:4328-4380): the into-wall part of a wet open cell's
// velocity goes, wall by wall.
__global__ void k_block_into_solid(float* __restrict__ u, float* __restrict__ v, const float* __restrict__ h,
                                   const uint8_t* __restrict__ solid, Region r, int res, float eps) {
  int i, j;
  if (!region_cell(r, &i, &j)) return;
  const size_t idx = (size_t)j * res + i;
  if (solid[idx] || h[idx] <= eps) return;
  float ux = u[idx], vx = v[idx];
  bool hit = false;
  for (int k = 0; k < 4; k++) {
    const int ni = i + kNb4[k][0], nj = j + kNb4[k][1];
    if (!grid_cell(ni, nj, res) || !solid[(size_t)nj * res + ni]) continue;
    const float nx = (float)kNb4[k][0], nz = (float)kNb4[k][1];
    const float vn = ux * nx + vx * nz;
    if (vn > 0.0f) {
      ux = ux - vn * nx;
      vx = vx - vn * nz;
      hit = true;
    }
  }
  if (hit) {
    u[idx] = ux;
    v[idx] = vx;
  }
}

// move_edge_bufs (:4178-4223) as a flux of one edge: no face next to a wall
// carries water; the dry-bank test always applies.
__device__ inline float wall_edge_dh(const uint8_t* solid, size_t d, size_t rcv, float hd, float hr, float remd,
                                     uint8_t gate_r, float bed_d, float bed_r, float speed, float dt, float cell) {
  if (solid[d] || solid[rcv]) return 0.0f;
  return edge_dh(hd, hr, remd, gate_r, true, bed_d, bed_r, speed, dt, cell);
}

// transport with solids (:3894-3977) for one cell: the serial edge loop's
// books of its two box-interior edges, as transport_cell keeps them.
__device__ inline float transport_cell_walled(const float* hin, const float* vel, const uint8_t* gate,
                                              const uint8_t* solid, const float* bed, size_t idx, size_t stride,
                                              int pos, int lo, int hi, float dt, float cell, bool gate_uniform) {
  auto gate_of = [&](size_t k) -> uint8_t { return gate_uniform ? (uint8_t)1 : gate[k]; };
  auto fl = [&](size_t k) -> float {
    const float speed = 0.5f * (vel[k] + vel[k + stride]);
    const float dh = wall_edge_dh(solid, k + stride, k, hin[k + stride], hin[k], 0.0f, gate_of(k), bed[k + stride],
                                  bed[k], speed, dt, cell);
    return speed > 0.0f ? 0.0f : dh;
  };
  auto fr = [&](size_t k, bool has_prev) -> float {
    const float speed = 0.5f * (vel[k] + vel[k + stride]);
    const float prev = has_prev ? fl(k - stride) : 0.0f;
    const float dh = wall_edge_dh(solid, k, k + stride, hin[k], hin[k + stride], 0.0f + prev, gate_of(k + stride),
                                  bed[k], bed[k + stride], speed, dt, cell);
    return speed > 0.0f ? dh : 0.0f;
  };
  const bool has_left = pos > lo, has_right = pos < hi;
  const float fl_prev = has_left ? fl(idx - stride) : 0.0f;
  const float fr_prev = has_left ? fr(idx - stride, pos - 1 > lo) : 0.0f;
  const float fl_here = has_right ? fl(idx) : 0.0f;
  const float fr_here = has_right ? fr(idx, has_left) : 0.0f;
  const float rem = (0.0f + fl_prev) + fr_here;
  const float add = (0.0f + fr_prev) + fl_here;
  const float hv = hin[idx];
  const float moved = fmaxf(hv - rem + add, 0.0f);
  return (rem > 0.0f || add > 0.0f) ? moved : hv;
}

__global__ void k_transport_walled(const float* __restrict__ hin, float* __restrict__ hout,
                                   const float* __restrict__ vel, const uint8_t* __restrict__ gate,
                                   const uint8_t* __restrict__ solid, const float* __restrict__ bed, Box b,
                                   int along_x, StepArgs a, float dt) {
  Region r{b.i0, b.i1, b.j0, b.j1};
  int i, j;
  const bool in = region_cell(r, &i, &j);
  double moved = 0.0;
  if (in) {
    const size_t idx = (size_t)j * a.res + i;
    if (along_x) {
      hout[idx] = transport_cell_walled(hin, vel, gate, solid, bed, idx, 1, i, b.i0, b.i1, dt, a.cell,
                                        a.gate_uniform);
    } else {
      const float was = hout[idx];
      const float now = transport_cell_walled(hin, vel, gate, solid, bed, idx, (size_t)a.res, j, b.j0, b.j1, dt,
                                              a.cell, a.gate_uniform);
      hout[idx] = now;
      moved = (double)fmaxf(now, 0.0f) - (double)fmaxf(was, 0.0f);
    }
  }
  if (!along_x && a.book) book_wave(&a.book->residue, moved);
}

// displace_occupied (:4382-4444), first half: a solid cell with water
// empties (its depth split over its open 4-neighbours, or, with none, over
// the open cells within two cells; mode 1 or 2) unless nothing nearby is
// open; shallower than 1e-8 it just dries.
__global__ void k_displace_give(float* __restrict__ h, float* __restrict__ u, float* __restrict__ v,
                                const uint8_t* __restrict__ solid, int res, float* __restrict__ share,
                                float* __restrict__ mode, SheetTotals* book) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  const bool in = i < res && j < res;
  double moved = 0.0;
  if (in) {
    const size_t idx = (size_t)j * res + i;
    float m = 0.0f, s = 0.0f;
    if (solid[idx]) {
      const float vol = h[idx];
      if (vol <= 1.0e-8f) {
        moved = -(double)fmaxf(vol, 0.0f);
        h[idx] = u[idx] = v[idx] = 0.0f;
      } else {
        int n4 = 0;
        for (int k = 0; k < 4; k++) {
          const int ni = i + kNb4[k][0], nj = j + kNb4[k][1];
          if (grid_cell(ni, nj, res) && !solid[(size_t)nj * res + ni]) n4++;
        }
        int n25 = 0;
        if (n4 == 0)
          for (int dj = -2; dj <= 2; dj++)
            for (int di = -2; di <= 2; di++) {
              const int ni = i + di, nj = j + dj;
              if (grid_cell(ni, nj, res) && !solid[(size_t)nj * res + ni]) n25++;
            }
        if (n4 > 0 || n25 > 0) {
          m = n4 > 0 ? 1.0f : 2.0f;
          s = vol / (float)(n4 > 0 ? n4 : n25);
          moved = -(double)vol;
          h[idx] = u[idx] = v[idx] = 0.0f;
        }
      }
    }
    share[idx] = s;
    mode[idx] = m;
  }
  if (book) book_wave(&book->residue, moved);
}

// Second half: an open cell adds the shares sent to it, sources in raster
// order, then h += give once.
__global__ void k_displace_take(float* __restrict__ h, const uint8_t* __restrict__ solid, int res,
                                const float* __restrict__ share, const float* __restrict__ mode, SheetTotals* book) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  const bool in = i < res && j < res;
  double moved = 0.0;
  if (in && !solid[(size_t)j * res + i]) {
    float give = 0.0f;
    for (int dj = -2; dj <= 2; dj++)
      for (int di = -2; di <= 2; di++) {
        const int si = i + di, sj = j + dj;
        if (!grid_cell(si, sj, res)) continue;
        const size_t s = (size_t)sj * res + si;
        const float m = mode[s];
        const bool four = (di == 0) != (dj == 0) && (di == 1 || di == -1 || dj == 1 || dj == -1);
        if ((m == 1.0f && four) || m == 2.0f) give = give + share[s];
      }
    if (give > 0.0f) {
      const size_t idx = (size_t)j * res + i;
      const float was = h[idx];
      const float now = was + give;
      h[idx] = now;
      moved = (double)now - (double)was;
    }
  }
  if (book) book_wave(&book->residue, moved);
}

inline dim3 grid_for(const Region& r, dim3 blk) {
  return dim3((unsigned)((r.i1 - r.i0 + 1 + (int)blk.x - 1) / (int)blk.x),
              (unsigned)((r.j1 - r.j0 + 1 + (int)blk.y - 1) / (int)blk.y));
}

// The in-place force pass of one box in raster Gauss-Seidel order: the
// kicks into (su, sv), then the confinement sweep; without confinement the
// kicks alone, in place.  Boxes over 252 rows take the global-memory sweep.
void gauss_seidel_force(const SheetPlanes& s, const Box& b, const StepArgs& a, hipStream_t st) {
  if (b.j1 - b.j0 + 1 + 4 > kRingRows) {
    hipLaunchKernelGGL(k_force_gauss_seidel, dim3(1), dim3(256), 0, st, s.h, s.u, s.v, s.nu, s.p, b, a);
    return;
  }
  Region box{b.i0, b.i1, b.j0, b.j1};
  dim3 blk(64, 4);
  dim3 g_box = grid_for(box, blk);
  if (a.vort > 0.0f) {
    hipLaunchKernelGGL(k_kick_box, g_box, blk, 0, st, s.h, s.u, s.v, s.nu, s.p, s.su, s.sv, box, a);
    hipLaunchKernelGGL(k_confine_gs_lds, dim3(1), dim3(256), 0, st, s.h, s.u, s.v, s.su, s.sv, b, a);
  } else {
    hipLaunchKernelGGL(k_kick_box, g_box, blk, 0, st, s.h, s.u, s.v, s.nu, s.p, s.u, s.v, box, a);
  }
}

}  // namespace

// One substep on one box (substep, src/fluid_core.rs:1879-2082).  do_flux 0
// stops after the damping (pace 3's pressure pass).
void exact_substep(const SheetPlanes& s, const Box& b, const StepArgs& a, bool do_flux, bool gauss_seidel,
                   hipStream_t st) {
  const int res = a.res;
  Region rim{b.i0 >= 2 ? b.i0 - 2 : 0, b.i1 + 2 < res - 1 ? b.i1 + 2 : res - 1, b.j0 >= 2 ? b.j0 - 2 : 0,
             b.j1 + 2 < res - 1 ? b.j1 + 2 : res - 1};
  Region box{b.i0, b.i1, b.j0, b.j1};
  dim3 blk(64, 4);
  dim3 g_rim = grid_for(rim, blk), g_box = grid_for(box, blk);
  hipLaunchKernelGGL(k_eta, g_rim, blk, 0, st, s.h, s.bed, s.eta, rim, res);
  if (!a.wind) {
    hipLaunchKernelGGL(k_xblur2, g_rim, blk, 0, st, s.eta, s.tg, s.ts, rim, b, res);
    hipLaunchKernelGGL(k_zblur_pressure, g_rim, blk, 0, st, s.tg, s.ts, s.eta, s.p, rim, b, res);
  }
  if (gauss_seidel) {
    gauss_seidel_force(s, b, a, st);
  } else {
    hipLaunchKernelGGL(k_force, g_rim, blk, 0, st, s.h, s.u, s.v, s.nu, s.p, s.su, s.sv, rim, b, a);
    if (a.vort > 0.0f)
      hipLaunchKernelGGL(k_confine_jacobi, g_box, blk, 0, st, s.h, s.u, s.v, s.su, s.sv, box, a);
  }
  if (a.project) {
    hipLaunchKernelGGL(k_project_source, g_rim, blk, 0, st, s.u, s.v, s.eta, s.nu, s.bsrc, rim, b, a);
    hipLaunchKernelGGL(k_xblur2, g_rim, blk, 0, st, s.bsrc, s.tg, s.ts, rim, b, res);
    hipLaunchKernelGGL(k_zblur_q, g_rim, blk, 0, st, s.tg, s.ts, s.p, rim, b, res);
  }
  hipLaunchKernelGGL(k_damp, g_box, blk, 0, st, s.h, s.u, s.v, s.nu, s.p, box, a);
  if (!do_flux) return;
  exact_transport(s, b, a, a.dt, st);
  hipLaunchKernelGGL(k_dry_cut, g_box, blk, 0, st, s.h, s.u, s.v, b, a);
}

// Transport along x then z with time step dt (the two transport_open calls of
// the substep, or one pace-3 slice).
void exact_transport(const SheetPlanes& s, const Box& b, const StepArgs& a, float dt, hipStream_t st) {
  Region box{b.i0, b.i1, b.j0, b.j1};
  dim3 blk(64, 4);
  dim3 g_box = grid_for(box, blk);
  hipLaunchKernelGGL(k_transport, g_box, blk, 0, st, s.h, s.h1, s.u, s.gate, s.bed, b, 1, a, dt);
  // Cells of the box now hold the x-moved depth in h1; copy the box back
  // through the z pass, which writes h.
  hipLaunchKernelGGL(k_transport, g_box, blk, 0, st, s.h1, s.h, s.v, s.gate, s.bed, b, 0, a, dt);
}

// One substep of one box beside walls (substep with solid = Some,
// src/fluid_core.rs:1879-2082).  par: the gravity kicks all first, then the
// confinement sweep (force_gravity_solid_par, the reference's form for a
// box of at least 48 x 32 cells when row splits are allowed); otherwise kick
// and confinement cell by cell.  Scratch after the damping: bsrc holds the
// redirect's removed depth, tg / ts / su / sv its four walls' shares, h1 the
// flank bits (then the transport's x pass).
void exact_substep_walled(const SheetPlanes& s, const Box& b, const StepArgs& a0, bool par, hipStream_t st) {
  StepArgs a = a0;
  a.project = 1;
  const int res = a.res;
  Region rim{b.i0 >= 2 ? b.i0 - 2 : 0, b.i1 + 2 < res - 1 ? b.i1 + 2 : res - 1, b.j0 >= 2 ? b.j0 - 2 : 0,
             b.j1 + 2 < res - 1 ? b.j1 + 2 : res - 1};
  Region rim1{b.i0 >= 1 ? b.i0 - 1 : 0, b.i1 + 1 < res - 1 ? b.i1 + 1 : res - 1, b.j0 >= 1 ? b.j0 - 1 : 0,
              b.j1 + 1 < res - 1 ? b.j1 + 1 : res - 1};
  Region box{b.i0, b.i1, b.j0, b.j1};
  dim3 blk(64, 4);
  dim3 g_rim = grid_for(rim, blk), g_rim1 = grid_for(rim1, blk), g_box = grid_for(box, blk);
  hipLaunchKernelGGL(k_eta, g_rim, blk, 0, st, s.h, s.bed, s.eta, rim, res);
  hipLaunchKernelGGL(k_lift_solid, g_box, blk, 0, st, s.eta, s.h, s.solid, box, res, a.eps);
  hipLaunchKernelGGL(k_xblur2, g_rim, blk, 0, st, s.eta, s.tg, s.ts, rim, b, res);
  hipLaunchKernelGGL(k_zblur_pressure, g_rim, blk, 0, st, s.tg, s.ts, s.eta, s.p, rim, b, res);
  hipLaunchKernelGGL(k_noslip_ghost, g_box, blk, 0, st, s.u, s.v, s.h, s.solid, box, res, a.eps);
  if (a.wind) {
    gauss_seidel_force(s, b, a, st);  // su, sv are free until the redirect
  } else if (par) {
    hipLaunchKernelGGL(k_kick_walled, g_box, blk, 0, st, s.h, s.u, s.v, s.nu, s.p, s.solid, box, a);
    if (a.vort > 0.0f)
      hipLaunchKernelGGL(k_sweep_walled, dim3(1), dim3(256), 0, st, s.h, s.u, s.v, s.nu, s.p, s.solid, b, a, 0);
  } else {
    hipLaunchKernelGGL(k_sweep_walled, dim3(1), dim3(256), 0, st, s.h, s.u, s.v, s.nu, s.p, s.solid, b, a, 1);
  }
  hipLaunchKernelGGL(k_project_source, g_rim, blk, 0, st, s.u, s.v, s.eta, s.nu, s.bsrc, rim, b, a);
  hipLaunchKernelGGL(k_xblur2, g_rim, blk, 0, st, s.bsrc, s.tg, s.ts, rim, b, res);
  hipLaunchKernelGGL(k_zblur_q, g_rim, blk, 0, st, s.tg, s.ts, s.p, rim, b, res);
  hipLaunchKernelGGL(k_damp, g_box, blk, 0, st, s.h, s.u, s.v, s.nu, s.p, box, a);
  hipLaunchKernelGGL(k_redirect_give, g_box, blk, 0, st, s.h, s.u, s.v, s.solid, s.bed, box, a, s.bsrc, s.tg, s.ts,
                     s.su, s.sv, s.h1);
  hipLaunchKernelGGL(k_redirect_take, g_rim1, blk, 0, st, s.h, rim1, b, a, s.bsrc, s.tg, s.ts, s.su, s.sv, s.h1);
  hipLaunchKernelGGL(k_block_into_solid, g_box, blk, 0, st, s.u, s.v, s.h, s.solid, box, res, a.eps);
  hipLaunchKernelGGL(k_noslip_ghost, g_box, blk, 0, st, s.u, s.v, s.h, s.solid, box, res, a.eps);
  hipLaunchKernelGGL(k_transport_walled, g_box, blk, 0, st, s.h, s.h1, s.u, s.gate, s.solid, s.bed, b, 1, a, a.dt);
  hipLaunchKernelGGL(k_transport_walled, g_box, blk, 0, st, s.h1, s.h, s.v, s.gate, s.solid, s.bed, b, 0, a, a.dt);
  hipLaunchKernelGGL(k_dry_cut, g_box, blk, 0, st, s.h, s.u, s.v, b, a);
}

// displace_occupied (:4382-4444) over the whole grid; share and mode are
// scratch planes of res * res floats.
void displace_occupied(float* h, float* u, float* v, const uint8_t* solid, int res, float* share, float* mode,
                       SheetTotals* book, hipStream_t st) {
  dim3 blk(64, 4);
  dim3 g((unsigned)((res + 63) / 64), (unsigned)((res + 3) / 4));
  hipLaunchKernelGGL(k_displace_give, g, blk, 0, st, h, u, v, solid, res, share, mode, book);
  hipLaunchKernelGGL(k_displace_take, g, blk, 0, st, h, solid, res, share, mode, book);
}

void exact_dry_cut(const SheetPlanes& s, const Box& b, const StepArgs& a, hipStream_t st) {
  Region box{b.i0, b.i1, b.j0, b.j1};
  dim3 blk(64, 4);
  hipLaunchKernelGGL(k_dry_cut, grid_for(box, blk), blk, 0, st, s.h, s.u, s.v, b, a);
}

}  // namespace silt
