// Data and inference kernels of the tire sub-grid network (surrogate.hpp).
#include "surrogate.hpp"

#include <math.h>

#include "mlp.hpp"
#include "sheet_fused.hpp"

namespace silt {

namespace {

// Nominal contact of a tire in the pressure channel: a 3.5 kN wheel load on
// a 0.30 m long contact of the tire's width, in units of 100 kPa.
constexpr float kContactLength = 0.30f;
constexpr float kWheelLoad = 3500.0f;

// Context cell a of region r for pushes p[0..n) with mean velocity (vtx, vtz).
// Context cell a of region r for pushes p[0..n) with mean velocity (vtx,
// vtz), computed by the 8 lanes sub = 0..7 of an aligned group: each lane
// tests one row of the 8 x 8 coverage samples of every push, the integer
// hit counts are summed over the group, and lane 0 forms the pressure in
// push order and stores the cell.
__device__ inline void sur_context_cell(const SurGrid& c, const SurRegion& r, const CorePush* pushes, int n, float vtx,
                                        float vtz, float* __restrict__ ctx, int a, int sub) {
  const bool cell_in = a < kSurCtx * kSurCtx;
  const int ca = a % kSurCtx, cb = a / kSurCtx;
  const int i = r.i0 - kSurReach + ca, j = r.j0 - kSurReach + cb;
  const size_t plane = (size_t)kSurCtx * kSurCtx;
  const bool on_grid = cell_in && i >= 0 && j >= 0 && i < c.res && j < c.res;
  float p = 0.0f;
  const float x0 = c.ox + (float)i * c.cell, z0 = c.oz + (float)j * c.cell;
  for (int t = 0; t < n; t++) {
    const CorePush q = pushes[t];
    const float fl = sqrtf(q.fx * q.fx + q.fz * q.fz);
    if (!(fl > 0.0f)) continue;
    const float fx = q.fx / fl, fz = q.fz / fl;
    // Coverage of the cell by the contact rectangle on 8 x 8 sample points
    // (2.5 cm apart at 0.2 m cells, about the 2 cm patch's spacing).
    int hits = 0;
    if (on_grid)
      for (int s = sub * 8; s < sub * 8 + 8; s++) {
        const float xs = x0 + ((float)(s & 7) + 0.5f) * (c.cell * 0.125f);
        const float zs = z0 + ((float)(s >> 3) + 0.5f) * (c.cell * 0.125f);
        const float dx = xs - q.x, dz = zs - q.z;
        const float along = dx * fx + dz * fz, across = dz * fx - dx * fz;
        hits += (fabsf(along) <= 0.5f * kContactLength && fabsf(across) <= 0.5f * q.width) ? 1 : 0;
      }
    hits += __shfl_xor(hits, 1, 8);
    hits += __shfl_xor(hits, 2, 8);
    hits += __shfl_xor(hits, 4, 8);
    if (hits) p += (kWheelLoad / (kContactLength * fmaxf(q.width, 0.05f))) * 1.0e-5f * ((float)hits / 64.0f);
  }
  if (!cell_in || sub != 0) return;
  float v[kSurChannels] = {0, 0, 0, 0, 0, 0, 0};
  if (on_grid) {
    const size_t idx = (size_t)j * c.res + i;
    const float hw = c.water[idx];
    const bool wet = hw > 1.0e-6f;
    v[0] = c.bed[idx];
    v[1] = hw * 10.0f;
    v[2] = c.slurry[idx] * 10.0f;
    v[3] = wet ? (c.uw[idx] - vtx) * 0.125f : 0.0f;
    v[4] = wet ? (c.vw[idx] - vtz) * 0.125f : 0.0f;
    v[5] = p;
    v[6] = 1.0f;
  }
  for (int ch = 0; ch < kSurChannels; ch++) ctx[(size_t)ch * plane + a] = v[ch];
}

// Eight lanes per context cell.
__global__ __launch_bounds__(256) void k_sur_context(SurGrid c, SurRegion r, PushList pl, float vtx, float vtz,
                                                     float* __restrict__ ctx) {
  const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  sur_context_cell(c, r, pl.p, pl.n, vtx, vtz, ctx, g >> 3, g & 7);
}

// One vehicle per blockIdx.y.
__global__ __launch_bounds__(256) void k_sur_context_many(SurGrid c, const TireJob* __restrict__ jobs,
                                                          const CorePush* __restrict__ pushes,
                                                          float* __restrict__ ctx) {
  const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const TireJob J = jobs[blockIdx.y];
  if (J.r.nx == 0) return;
  sur_context_cell(c, J.r, pushes + J.push.p0, J.push.np, J.vtx, J.vtz, ctx + (size_t)blockIdx.y * kSurCtxFloats,
                   g >> 3, g & 7);
}

// The square's symmetries as x' = M x, M = {m00, m01, m10, m11}.
__constant__ int kSym[8][4] = {{1, 0, 0, 1},  {0, -1, 1, 0}, {-1, 0, 0, -1}, {0, 1, -1, 0},
                               {-1, 0, 0, 1}, {1, 0, 0, -1}, {0, 1, 1, 0},   {0, -1, -1, 0}};

// Input k of cell (ci, cj) of a region, under symmetry t: channel k / 256,
// stencil row (k % 256) / 16, column k % 16 (offsets -8..+7 in the turned
// frame, read at M^T of them, which stays within -8..+8).
__device__ inline float sur_input(const float* __restrict__ ctx, int ci, int cj, int k, int t) {
  const size_t plane = (size_t)kSurCtx * kSurCtx;
  const int ch = k >> 8, oz = ((k >> 4) & 15) - kSurReach, ox = (k & 15) - kSurReach;
  const int* m = kSym[t];
  // M^T (the inverse) of the turned offset.
  const int dx = m[0] * ox + m[2] * oz, dz = m[1] * ox + m[3] * oz;
  const size_t pos = (size_t)(cj + kSurReach + dz) * kSurCtx + (ci + kSurReach + dx);
  if (ch == 0) {
    if (ctx[6 * plane + pos] == 0.0f) return 0.0f;
    const size_t centre = (size_t)(cj + kSurReach) * kSurCtx + (ci + kSurReach);
    return (ctx[pos] - ctx[centre]) * 10.0f;
  }
  if (ch == 3 || ch == 4) {
    const float u = ctx[3 * plane + pos], v = ctx[4 * plane + pos];
    return ch == 3 ? m[0] * u + m[1] * v : m[2] * u + m[3] * v;
  }
  return ctx[(size_t)ch * plane + pos];
}

// Edge of the unturned frame that edge e (0 +x, 1 -x, 2 +z, 3 -z) of the
// turned frame came from.
__device__ inline int sur_edge_from(int e, int t) {
  const int* m = kSym[t];
  const int ex = e == 0 ? 1 : e == 1 ? -1 : 0, ez = e == 2 ? 1 : e == 3 ? -1 : 0;
  const int x = m[0] * ex + m[2] * ez, z = m[1] * ex + m[3] * ez;
  return x == 1 ? 0 : x == -1 ? 1 : z == 1 ? 2 : 3;
}

__global__ void k_sur_gather_region(const float* __restrict__ ctx, SurRegion r, uint16_t* __restrict__ xb) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t cells = (size_t)r.nx * r.nz;
  if (t >= cells * kMlpIn) return;
  const int q = (int)(t / kMlpIn), k = (int)(t % kMlpIn);
  xb[t] = f2bf(sur_input(ctx, q % r.nx, q / r.nx, k, 0));
}

__global__ void k_sur_gather_many(const float* __restrict__ ctx, const TireJob* __restrict__ jobs,
                                  uint16_t* __restrict__ xb) {
  const TireJob J = jobs[blockIdx.y];
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t cells = (size_t)J.r.nx * J.r.nz;
  if (t >= cells * kMlpIn) return;
  const int q = (int)(t / kMlpIn), k = (int)(t % kMlpIn);
  xb[(size_t)J.row * kMlpIn + t] =
      f2bf(sur_input(ctx + (size_t)blockIdx.y * kSurCtxFloats, q % J.r.nx, q / J.r.nx, k, 0));
}

__device__ inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__global__ void k_sur_gather_random(const float* __restrict__ ctx, const float* __restrict__ tgt,
                                    const SurRegion* __restrict__ regions, int nsteps, uint32_t seed, int batch,
                                    uint16_t* __restrict__ xb, float* __restrict__ tg) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)batch * kMlpIn) return;
  const int b = (int)(t / kMlpIn), k = (int)(t % kMlpIn);
  const uint32_t h1 = mix32(seed * 0x9e3779b9u + (uint32_t)b);
  const uint32_t h2 = mix32(h1 ^ 0x85ebca6bu);
  const int s = (int)(h1 % (uint32_t)nsteps);
  const SurRegion r = regions[s];
  const int q = (int)(h2 % (uint32_t)(r.nx * r.nz));
  const int sym = (int)(mix32(h2 ^ 0x27d4eb2fu) & 7u);
  const float* c = ctx + (size_t)s * kSurCtxFloats;
  xb[t] = f2bf(sur_input(c, q % r.nx, q / r.nx, k, sym));
  if (k < kMlpOut) {
    const int layer = k >> 2, e = k & 3;
    tg[(size_t)b * kMlpOut + k] = tgt[(size_t)s * kSurTgtFloats + (size_t)q * 8 + layer * 4 + sur_edge_from(e, sym)];
  }
}

// Least-norm edge fluxes of a depth difference over the region: solve
// L phi = d with the Neumann graph Laplacian of the nx x nz grid by its
// cosine eigenbasis (cos(pi k (i + 1/2) / n), eigenvalue -(2 - 2 cos(pi k / n))
// per axis), then the outflow from a to a neighbour b is phi_a - phi_b.
__global__ __launch_bounds__(256) void k_sur_targets(const float* __restrict__ hn_w, const float* __restrict__ hc_w,
                                                     const float* __restrict__ hn_s, const float* __restrict__ hc_s,
                                                     int res, SurRegion r, float* __restrict__ tgt) {
  __shared__ double d[kSurMax][kSurMax];
  __shared__ double t1[kSurMax][kSurMax];
  __shared__ double mean;
  const int layer = (int)blockIdx.x;
  const float* hn = layer ? hn_s : hn_w;
  const float* hc = layer ? hc_s : hc_w;
  const int nx = r.nx, nz = r.nz, cells = nx * nz, tid = (int)threadIdx.x;
  const double pi = 3.14159265358979323846;
  for (int e = tid; e < cells; e += 256) {
    const int i = e % nx, j = e / nx;
    const size_t idx = (size_t)(r.j0 + j) * res + (r.i0 + i);
    d[j][i] = (double)hn[idx] - (double)hc[idx];
  }
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int j = 0; j < nz; j++)
      for (int i = 0; i < nx; i++) s += d[j][i];
    mean = s / (double)cells;
  }
  __syncthreads();
  for (int e = tid; e < cells; e += 256) d[e / nx][e % nx] -= mean;
  __syncthreads();
  // t1[j][k] = sum_i d[j][i] cx(k, i)
  for (int e = tid; e < cells; e += 256) {
    const int k = e % nx, j = e / nx;
    double s = 0.0;
    for (int i = 0; i < nx; i++) s += d[j][i] * cos(pi * k * (i + 0.5) / nx);
    t1[j][k] = s;
  }
  __syncthreads();
  // d[l][k] = sum_j t1[j][k] cz(l, j) / lambda(k, l)
  for (int e = tid; e < cells; e += 256) {
    const int k = e % nx, l = e / nx;
    double s = 0.0;
    for (int j = 0; j < nz; j++) s += t1[j][k] * cos(pi * l * (j + 0.5) / nz);
    const double lam = -(2.0 - 2.0 * cos(pi * k / nx)) - (2.0 - 2.0 * cos(pi * l / nz));
    d[l][k] = (k == 0 && l == 0) ? 0.0 : s / lam;
  }
  __syncthreads();
  // t1[l][i] = sum_k d[l][k] w(k) cx(k, i)
  for (int e = tid; e < cells; e += 256) {
    const int i = e % nx, l = e / nx;
    double s = 0.0;
    for (int k = 0; k < nx; k++) s += d[l][k] * (k ? 2.0 / nx : 1.0 / nx) * cos(pi * k * (i + 0.5) / nx);
    t1[l][i] = s;
  }
  __syncthreads();
  // phi[j][i] = sum_l t1[l][i] w(l) cz(l, j), into d
  for (int e = tid; e < cells; e += 256) {
    const int i = e % nx, j = e / nx;
    double s = 0.0;
    for (int l = 0; l < nz; l++) s += t1[l][i] * (l ? 2.0 / nz : 1.0 / nz) * cos(pi * l * (j + 0.5) / nz);
    d[j][i] = s;  // safe: each (j, i) is read only from t1 here
  }
  __syncthreads();
  for (int e = tid; e < cells; e += 256) {
    const int i = e % nx, j = e / nx;
    const double p = d[j][i];
    float* o = tgt + (size_t)e * 8 + layer * 4;
    o[0] = i + 1 < nx ? (float)((p - d[j][i + 1]) * kSurTargetScale) : 0.0f;  // E
    o[1] = i > 0 ? (float)((p - d[j][i - 1]) * kSurTargetScale) : 0.0f;       // W
    o[2] = j + 1 < nz ? (float)((p - d[j + 1][i]) * kSurTargetScale) : 0.0f;  // N
    o[3] = j > 0 ? (float)((p - d[j - 1][i]) * kSurTargetScale) : 0.0f;       // S
  }
}

// Outflow from cell q to its neighbour in direction dir (0 E, 1 W, 2 N, 3 S)
// of one layer, in metres; zero through the region's border.
__device__ inline float sur_face(const float* __restrict__ y, const SurRegion& r, int q, int dir, int layer) {
  const int i = q % r.nx, j = q / r.nx;
  int nb;
  if (dir == 0) {
    if (i + 1 >= r.nx) return 0.0f;
    nb = q + 1;
  } else if (dir == 1) {
    if (i == 0) return 0.0f;
    nb = q - 1;
  } else if (dir == 2) {
    if (j + 1 >= r.nz) return 0.0f;
    nb = q + r.nx;
  } else {
    if (j == 0) return 0.0f;
    nb = q - r.nx;
  }
  const int opp = dir ^ 1;
  const float a = y[(size_t)q * 8 + layer * 4 + dir], b = y[(size_t)nb * 8 + layer * 4 + opp];
  return 0.5f * (a - b) * (1.0f / kSurTargetScale);
}

__device__ inline int sur_nb(const SurRegion& r, int q, int dir) {
  return dir == 0 ? q + 1 : dir == 1 ? q - 1 : dir == 2 ? q + r.nx : q - r.nx;
}

// Factor and apply in one workgroup per region (2 x 20 x 20 lanes at most):
// the donor factors of both layers go to LDS, one barrier, then every cell
// applies its faces: a donor whose predicted outflow exceeds its depth gives
// all of its outflows scaled down to what it holds.
// With totals, each cell's residue ((new - old) + what its faces carried
// out, in FP64; the faces' FP32 fluxes cancel in pairs) is summed per wave
// and layer and added to that layer's residue line.
__device__ inline void sur_factor_apply_wg(float* __restrict__ water, float* __restrict__ slurry, int res,
                                           const float* __restrict__ yv, const SurRegion& r, float* fac,
                                           SheetTotals* totals, SheetStats* stats) {
  const int t = (int)threadIdx.x;
  const int cells = r.nx * r.nz;
  const bool mine = t < 2 * cells;
  const int layer = mine ? t / cells : 0, q = mine ? t % cells : 0;
  const size_t idx = (size_t)(r.j0 + q / r.nx) * res + (r.i0 + q % r.nx);
  float* h = layer ? slurry : water;
  if (mine) {
    const float hv = h[idx];
    float out = 0.0f;
    for (int dir = 0; dir < 4; dir++) out += fmaxf(sur_face(yv, r, q, dir, layer), 0.0f);
    fac[t] = out > hv ? 0.999999f * hv / out : 1.0f;
  }
  __syncthreads();
  double res_w = 0.0, res_s = 0.0, dv_w = 0.0, dv_s = 0.0;
  if (mine) {
    float net = 0.0f;
    double out = 0.0;
    for (int dir = 0; dir < 4; dir++) {
      const float f = sur_face(yv, r, q, dir, layer);
      if (f == 0.0f) continue;
      const int nb = sur_nb(r, q, dir);
      const float moved = f * (f > 0.0f ? fac[layer * cells + q] : fac[layer * cells + nb]);
      net += moved;
      out += (double)moved;
    }
    const float hv = h[idx];
    const float hn = hv - net;
    h[idx] = hn;
    const double rr = ((double)hn - (double)hv) + out;
    res_w = layer == 0 ? rr : 0.0;
    res_s = layer == 1 ? rr : 0.0;
    // The statistics' volume is the FP64 sum of max(h, 0): its change here.
    const double dv = (double)fmaxf(hn, 0.0f) - (double)fmaxf(hv, 0.0f);
    dv_w = layer == 0 ? dv : 0.0;
    dv_s = layer == 1 ? dv : 0.0;
  }
  if (!totals && !stats) return;
  for (int off = 32; off > 0; off >>= 1) {
    res_w += __shfl_xor(res_w, off);
    res_s += __shfl_xor(res_s, off);
    dv_w += __shfl_xor(dv_w, off);
    dv_s += __shfl_xor(dv_s, off);
  }
  if ((t & 63) == 0) {
    if (totals && res_w != 0.0) atomicAdd(&totals[0].residue, res_w);
    if (totals && res_s != 0.0) atomicAdd(&totals[1].residue, res_s);
    if (stats && dv_w != 0.0) atomicAdd(&stats[0].vol, dv_w);
    if (stats && dv_s != 0.0) atomicAdd(&stats[1].vol, dv_s);
  }
}

__global__ __launch_bounds__(1024) void k_sur_factor_apply(float* __restrict__ water, float* __restrict__ slurry,
                                                           int res, const float* __restrict__ y, SurRegion r,
                                                           SheetTotals* totals, SheetStats* stats) {
  __shared__ float fac[2 * kSurMax * kSurMax];
  sur_factor_apply_wg(water, slurry, res, y, r, fac, totals, stats);
}

__global__ __launch_bounds__(1024) void k_sur_factor_apply_many(float* __restrict__ water, float* __restrict__ slurry,
                                                                int res, const float* __restrict__ y,
                                                                const TireJob* __restrict__ jobs,
                                                                SheetTotals* totals, SheetStats* stats) {
  __shared__ float fac[2 * kSurMax * kSurMax];
  const TireJob J = jobs[blockIdx.x];
  sur_factor_apply_wg(water, slurry, res, y + (size_t)J.row * kMlpOut, J.r, fac, totals, stats);
}

__global__ void k_sur_max_abs_diff(const float* __restrict__ a, const float* __restrict__ b, int n,
                                   unsigned* out) {
  float m = 0.0f;
  for (int k = (int)(blockIdx.x * blockDim.x + threadIdx.x); k < n; k += (int)(gridDim.x * blockDim.x))
    m = fmaxf(m, fabsf(a[k] - b[k]));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

__global__ void k_sur_region_sq_diff(const float* __restrict__ a, const float* __restrict__ b, int res, SurRegion r,
                                     double* out) {
  const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  double s = 0.0;
  if (q < r.nx * r.nz) {
    const size_t idx = (size_t)(r.j0 + q / r.nx) * res + (r.i0 + q % r.nx);
    const double d = (double)a[idx] - (double)b[idx];
    s = d * d;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0 && s != 0.0) atomicAdd(out, s);
}

}  // namespace

void sur_region_sq_diff(const float* a, const float* b, int res, const SurRegion& r, double* out, hipStream_t st) {
  const int n = r.nx * r.nz;
  hipLaunchKernelGGL(k_sur_region_sq_diff, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, res, r, out);
}

SurRegion sur_region_at(const CorePush* p, int n, float ox, float oz, float cell, int res) {
  SurRegion r{0, 0, 0, 0};
  if (n <= 0) return r;
  float cx = 0.0f, cz = 0.0f;
  for (int k = 0; k < n; k++) {
    cx += p[k].x;
    cz += p[k].z;
  }
  cx /= (float)n;
  cz /= (float)n;
  const float extent = (float)kFineRes * kFineCell;
  const float fx = floorf((cx - extent * 0.5f) / cell) * cell;
  const float fz = floorf((cz - extent * 0.5f) / cell) * cell;
  const int pi0 = (int)lroundf((fx - ox) / cell) - 2;
  const int pj0 = (int)lroundf((fz - oz) / cell) - 2;
  const int i0 = pi0 > 0 ? pi0 : 0, j0 = pj0 > 0 ? pj0 : 0;
  const int i1 = pi0 + kSurMax < res ? pi0 + kSurMax : res, j1 = pj0 + kSurMax < res ? pj0 + kSurMax : res;
  if (i1 - i0 < 2 || j1 - j0 < 2) return r;
  return SurRegion{i0, j0, i1 - i0, j1 - j0};
}

void sur_context(const SurGrid& c, const SurRegion& r, const PushList& pl, float* ctx, hipStream_t st) {
  float vx, vz;
  sur_mean_velocity(pl.p, pl.n, &vx, &vz);
  const int n = kSurCtx * kSurCtx * 8;
  hipLaunchKernelGGL(k_sur_context, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, c, r, pl, vx, vz, ctx);
}

void sur_mean_velocity(const CorePush* p, int n, float* vx, float* vz) {
  float ax = 0.0f, az = 0.0f;
  for (int t = 0; t < n; t++) {
    const float fl = sqrtf(p[t].fx * p[t].fx + p[t].fz * p[t].fz);
    if (fl > 0.0f) {
      ax += p[t].speed * p[t].fx / fl;
      az += p[t].speed * p[t].fz / fl;
    }
  }
  if (n) {
    ax /= (float)n;
    az /= (float)n;
  }
  *vx = ax;
  *vz = az;
}

void sur_context_many(const SurGrid& g, const TireJob* jobs, const CorePush* pushes, int n, float* ctx,
                      hipStream_t st) {
  if (n <= 0) return;
  const int lanes = kSurCtx * kSurCtx * 8;
  hipLaunchKernelGGL(k_sur_context_many, dim3((unsigned)((lanes + 255) / 256), (unsigned)n), dim3(256), 0, st, g, jobs,
                     pushes, ctx);
}

void sur_gather_many(const float* ctx, const TireJob* jobs, int n, uint16_t* xb, hipStream_t st) {
  if (n <= 0) return;
  const size_t most = (size_t)kSurMax * kSurMax * kMlpIn;
  hipLaunchKernelGGL(k_sur_gather_many, dim3((unsigned)((most + 255) / 256), (unsigned)n), dim3(256), 0, st, ctx,
                     jobs, xb);
}

void sur_apply_many(float* water, float* slurry, int res, const float* y, const TireJob* jobs, int n, float* fac,
                    hipStream_t st, SheetTotals* totals, SheetStats* stats) {
  (void)fac;  // the factors live in LDS
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sur_factor_apply_many, dim3((unsigned)n), dim3(1024), 0, st, water, slurry, res, y, jobs,
                     totals, stats);
}

void sur_gather_region(const float* ctx, const SurRegion& r, uint16_t* xb, hipStream_t st) {
  const size_t n = (size_t)r.nx * r.nz * kMlpIn;
  hipLaunchKernelGGL(k_sur_gather_region, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ctx, r, xb);
}

void sur_gather_random(const float* ctx, const float* tgt, const SurRegion* regions, int nsteps, uint32_t seed,
                       int batch, uint16_t* xb, float* tg, hipStream_t st) {
  const size_t n = (size_t)batch * kMlpIn;
  hipLaunchKernelGGL(k_sur_gather_random, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ctx, tgt, regions,
                     nsteps, seed, batch, xb, tg);
}

void sur_targets(const float* hn_w, const float* hc_w, const float* hn_s, const float* hc_s, int res,
                 const SurRegion& r, float* tgt, hipStream_t st) {
  hipLaunchKernelGGL(k_sur_targets, dim3(2), dim3(256), 0, st, hn_w, hc_w, hn_s, hc_s, res, r, tgt);
}

void sur_apply(float* water, float* slurry, int res, const float* y, const SurRegion& r, float* fac,
               hipStream_t st, SheetTotals* totals, SheetStats* stats) {
  (void)fac;  // the factors live in LDS
  if (r.nx <= 0 || r.nz <= 0) return;
  hipLaunchKernelGGL(k_sur_factor_apply, dim3(1), dim3(1024), 0, st, water, slurry, res, y, r, totals, stats);
}

void sur_max_abs_diff(const float* a, const float* b, int n, unsigned* out, hipStream_t st) {
  hipLaunchKernelGGL(k_sur_max_abs_diff, dim3(64), dim3(256), 0, st, a, b, n, out);
}

}  // namespace silt
