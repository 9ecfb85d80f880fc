// Kernels of the two-layer FluidCore step on the sheet path (the 40 m coarse
// window and the 2 cm patch under the tires): tire shoves, spray peel, the
// soft window edge, the slurry viscosity field, the coarse -> fine copy and
// the fine -> coarse sum, the near-field residual tiles, and the publish maps.
// Each kernel restates one reference function (cited at the kernel) with the
// same per-cell operation order, so a cell's arithmetic matches the
// reference bit for bit; only the order of independent cells changes.
#include "core.hpp"

namespace silt {

namespace {

__device__ inline int span_lo(float origin, float cell, int res, float center, float reach) {
  float lo = fmaxf(floorf((center - reach - origin) / cell - 1.0f), 0.0f);
  int l = (int)lo;
  return l < res ? l : res;
}

__device__ inline int span_hi(float origin, float cell, int res, float center, float reach, int lo) {
  float hi = fmaxf(ceilf((center + reach - origin) / cell + 2.0f), 0.0f);
  int h = (int)hi;
  h = h < res ? h : res;
  return h > lo ? h : lo;
}

// push_on (src/fluid_core.rs:8169-8239): pushes p[0..n), in order, on cell
// (i, j).
__device__ inline void push_on_cell(const GridGeo& g, float* water, float* slurry, float* uw, float* vw, float* us,
                                    float* vs, const CorePush* pushes, int n, int i, int j) {
  size_t idx = (size_t)j * g.res + i;
  const float cell = g.cell;
  for (int q = 0; q < n; q++) {
    const CorePush& p = pushes[q];
    float len = fmaxf(sqrtf(p.fx * p.fx + p.fz * p.fz), 1.0e-4f);
    float fx = p.fx / len, fz = p.fz / len;
    float rx = fz, rz = -fx;
    float speed = fmaxf(p.speed, 0.0f);
    float reach = hypotf(p.radius * 1.4f, p.width * 0.65f + cell) + cell;
    int a0 = span_lo(g.ox, cell, g.res, p.x, reach), a1 = span_hi(g.ox, cell, g.res, p.x, reach, a0);
    int b0 = span_lo(g.oz, cell, g.res, p.z, reach), b1 = span_hi(g.oz, cell, g.res, p.z, reach, b0);
    if (i < a0 || i >= a1 || j < b0 || j >= b1) continue;
    float x = g.ox + ((float)i + 0.5f) * cell;
    float z = g.oz + ((float)j + 0.5f) * cell;
    float dx = x - p.x, dz = z - p.z;
    float along = dx * fx + dz * fz;
    float lat = dx * rx + dz * rz;
    if (fabsf(lat) > p.width * 0.65f + cell) continue;
    if (along < -p.radius * 1.4f || along > p.radius * 1.1f) continue;
    float flank = clampf(fabsf(lat) / (p.width * 0.65f + cell), 0.0f, 1.0f);
    float shove = fmaxf(speed, 0.4f) * (1.0f - flank * flank);
    float sg = copysignf(1.0f, lat);
    if (water[idx] > 0.004f) {
      uw[idx] = uw[idx] + fx * shove * 0.85f;
      vw[idx] = vw[idx] + fz * shove * 0.85f;
      uw[idx] = uw[idx] + rx * sg * shove * 0.18f * flank;
      vw[idx] = vw[idx] + rz * sg * shove * 0.18f * flank;
    }
    if (slurry[idx] > 0.004f) {
      float fore = along > 0.0f ? 1.0f : 0.72f;
      us[idx] = us[idx] + fx * shove * fore;
      vs[idx] = vs[idx] + fz * shove * fore;
      us[idx] = us[idx] + rx * sg * shove * 0.28f * flank;
      vs[idx] = vs[idx] + rz * sg * shove * 0.28f * flank;
    }
  }
}

__global__ void k_push_on(GridGeo g, float* water, float* slurry, float* uw, float* vw, float* us, float* vs,
                          PushList pl, int i0, int j0, int w, int h) {
  int i = i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  int j = j0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i >= i0 + w || j >= j0 + h) return;
  push_on_cell(g, water, slurry, uw, vw, us, vs, pl.p, pl.n, i, j);
}

// Many groups of pushes on disjoint boxes, one group per blockIdx.z.
__global__ void k_push_on_many(GridGeo g, float* water, float* slurry, float* uw, float* vw, float* us, float* vs,
                               const PushBox* __restrict__ boxes, const CorePush* __restrict__ pushes) {
  const PushBox b = boxes[blockIdx.z];
  int i = b.i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  int j = b.j0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i >= b.i0 + b.w || j >= b.j0 + b.h) return;
  push_on_cell(g, water, slurry, uw, vw, us, vs, pushes + b.p0, b.np, i, j);
}

// peel_spray (src/fluid_core.rs:8241-8314) on one cell, every push in order.
// Events go to a buffer in arbitrary order; the peeled volume is summed in
// FP64 (the reference sums it in f32 in event order).
__global__ void k_peel(GridGeo g, float* water, const float* uw, const float* vw, PushList pl, int i0, int j0, int w,
                       int h, const uint8_t* hold, GridGeo cg, int skip_held, SprayEvent* ev, unsigned* nev,
                       unsigned cap, double* vol, double ledger_area) {
  int i = i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  int j = j0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i >= i0 + w || j >= j0 + h) return;
  size_t idx = (size_t)j * g.res + i;
  const float cell = g.cell, area = cell * cell;
  float x = g.ox + ((float)i + 0.5f) * cell;
  float z = g.oz + ((float)j + 0.5f) * cell;
  if (skip_held) {
    // coarse_held (src/fluid_core.rs:9664-9672)
    float ci = floorf((x - cg.ox) / 0.2f), cj = floorf((z - cg.oz) / 0.2f);
    if (ci >= 0.0f && cj >= 0.0f && ci < (float)cg.res && cj < (float)cg.res &&
        hold[(size_t)cj * cg.res + (size_t)ci] == 1)
      return;
  }
  for (int q = 0; q < pl.n; q++) {
    const CorePush& p = pl.p[q];
    if (p.speed < 0.8f) continue;
    float reach = p.radius * 1.3f + cell;
    int a0 = span_lo(g.ox, cell, g.res, p.x, reach), a1 = span_hi(g.ox, cell, g.res, p.x, reach, a0);
    int b0 = span_lo(g.oz, cell, g.res, p.z, reach), b1 = span_hi(g.oz, cell, g.res, p.z, reach, b0);
    if (i < a0 || i >= a1 || j < b0 || j >= b1) continue;
    float dx = x - p.x, dz = z - p.z;
    if (dx * dx + dz * dz > (p.radius * 1.3f) * (p.radius * 1.3f)) continue;
    float wv = water[idx];
    if (wv < 0.01f) continue;
    float take = fminf(wv * 0.04f, 0.004f);
    if (take * area < 1.0e-6f) continue;
    const float left = wv - take;
    water[idx] = left;
    // The ledger books the depth the cell lost, exactly.
    atomicAdd(vol, ((double)wv - (double)left) * ledger_area);
    unsigned k = atomicAdd(nev, 1u);
    if (k < cap) {
      SprayEvent e;
      e.x = x;
      e.y = 0.35f;
      e.z = z;
      e.vx = uw[idx] * 0.3f;
      e.vy = 1.4f + p.speed * 0.25f;
      e.vz = vw[idx] * 0.3f;
      e.volume = take * area;
      e.kind = 0;
      ev[k] = e;
    }
  }
}

// soften_border (src/fluid_core.rs:9594-9620), no open reach.
__global__ void k_soften(float* u, float* v, int res) {
  int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i >= res || j >= res) return;
  int edge = res / 12 > 2 ? res / 12 : 2;
  int di = min(i, res - 1 - i), dj = min(j, res - 1 - j);
  int d = min(di, dj);
  if (d >= edge) return;
  float t = (float)d / (float)edge;
  float w = t * t * (3.0f - 2.0f * t);
  size_t idx = (size_t)j * res + i;
  u[idx] = u[idx] * w;
  v[idx] = v[idx] * w;
}

// The slurry viscosity plane from grain and moisture (step_pair :8090-8098).
__global__ void k_nu(float* nu, const uint8_t* grain, const float* moisture, int n) {
  int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= n) return;
  nu[k] = dynamic_viscosity(grain[k], moisture[k]);
}

// Coarse -> fine copy of refine (src/fluid_core.rs:8418-8444), one fine cell.
__global__ void k_refine(CorePlanes c, FinePlanes f, const uint8_t* hold, FineAxes ax) {
  int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i >= kFineRes || j >= kFineRes) return;
  int ci = ax.col[i], cj = ax.row[j];
  if (ci < 0 || cj < 0) return;
  size_t fidx = (size_t)j * kFineRes + i, cidx = (size_t)cj * kWindowRes + ci;
  f.moisture[fidx] = c.moisture[cidx];
  f.grain[fidx] = c.grain[cidx];
  if (hold[cidx] != 1) {
    f.bed[fidx] = c.bed[cidx] + c.water[cidx];
    f.bed_s[fidx] = c.bed[cidx] + c.slurry[cidx];
    return;
  }
  f.bed[fidx] = c.bed[cidx];
  f.bed_s[fidx] = c.bed[cidx];
  f.water[fidx] = c.water[cidx];
  f.slurry[fidx] = c.slurry[cidx];
  f.uw[fidx] = c.uw[cidx];
  f.vw[fidx] = c.vw[cidx];
  f.us[fidx] = c.us[cidx];
  f.vs[fidx] = c.vs[cidx];
}

__global__ void k_zero_held(CorePlanes c, const uint8_t* hold, int ci0, int cj0, int w, int h) {
  int i = ci0 + (int)(blockIdx.x * blockDim.x + threadIdx.x), j = cj0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i >= ci0 + w || j >= cj0 + h) return;
  size_t idx = (size_t)j * kWindowRes + i;
  if (hold[idx] != 1) return;
  c.water[idx] = 0.0f;
  c.slurry[idx] = 0.0f;
  c.uw[idx] = 0.0f;
  c.vw[idx] = 0.0f;
  c.us[idx] = 0.0f;
  c.vs[idx] = 0.0f;
}

// Fine -> coarse sum of restrict_fine (src/fluid_core.rs:8462-8549) for one
// coarse cell: its fine cells in the fine raster order, then the hand-back.
// With the ledger: each coarse cell's FP64 gain less the FP64 volume of its
// fine cells goes to handoff[0..1].
__global__ void k_restrict(CorePlanes c, FinePlanes f, const uint8_t* hold, CoarseSpans sp, double* handoff) {
  int a = (int)(blockIdx.x * blockDim.x + threadIdx.x), b = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (a >= sp.ncol || b >= sp.nrow) return;
  int ci = sp.ci0 + a, cj = sp.cj0 + b;
  int i0 = sp.col_start[a], ni = sp.col_count[a];
  int j0 = sp.row_start[b], nj = sp.row_count[b];
  if (ni == 0 || nj == 0) return;
  const float fine_area = kFineCell * kFineCell, coarse_area = kCoarseCell * kCoarseCell;
  float w = 0.0f, s = 0.0f, auw = 0.0f, avw = 0.0f, aus = 0.0f, avs = 0.0f;
  double fine_w = 0.0, fine_s = 0.0;
  for (int j = j0; j < j0 + nj; j++)
    for (int i = i0; i < i0 + ni; i++) {
      size_t fidx = (size_t)j * kFineRes + i;
      float fw = f.water[fidx], fs = f.slurry[fidx];
      if (fw == 0.0f && fs == 0.0f) continue;
      fine_w += (double)fmaxf(fw, 0.0f);
      fine_s += (double)fmaxf(fs, 0.0f);
      w = w + fw * fine_area;
      s = s + fs * fine_area;
      auw = auw + f.uw[fidx] * fw * fine_area;
      avw = avw + f.vw[fidx] * fw * fine_area;
      aus = aus + f.us[fidx] * fs * fine_area;
      avs = avs + f.vs[fidx] * fs * fine_area;
    }
  size_t idx = (size_t)cj * kWindowRes + ci;
  bool held = hold[idx] == 1;
  float wd = w / coarse_area, sd = s / coarse_area;
  const float was_w = c.water[idx], was_s = c.slurry[idx];
  // Books the cell's gain beyond its fine cells' volume once it is written.
  auto book = [&](float now_w, float now_s) {
    if (!handoff) return;
    const double dw = ((double)fmaxf(now_w, 0.0f) - (double)fmaxf(was_w, 0.0f)) * kCoarseArea - fine_w * kFineArea;
    const double ds = ((double)fmaxf(now_s, 0.0f) - (double)fmaxf(was_s, 0.0f)) * kCoarseArea - fine_s * kFineArea;
    if (dw != 0.0) atomicAdd(&handoff[0], dw);
    if (ds != 0.0) atomicAdd(&handoff[1], ds);
  };
  if (!held && ni * nj >= 100) {
    if (wd <= 0.0f && sd <= 0.0f) {
      book(was_w, was_s);
      return;
    }
    float old_w = was_w, old_s = was_s;
    float new_w = old_w + wd, new_s = old_s + sd;
    if (new_w > 1.0e-6f) {
      c.uw[idx] = (old_w * c.uw[idx] * coarse_area + auw) / (new_w * coarse_area);
      c.vw[idx] = (old_w * c.vw[idx] * coarse_area + avw) / (new_w * coarse_area);
    }
    if (new_s > 1.0e-6f) {
      c.us[idx] = (old_s * c.us[idx] * coarse_area + aus) / (new_s * coarse_area);
      c.vs[idx] = (old_s * c.vs[idx] * coarse_area + avs) / (new_s * coarse_area);
    }
    c.water[idx] = new_w;
    c.slurry[idx] = new_s;
    book(new_w, new_s);
    return;
  }
  c.water[idx] = wd;
  c.slurry[idx] = sd;
  book(wd, sd);
  if (wd > 1.0e-6f) {
    c.uw[idx] = auw / (wd * coarse_area);
    c.vw[idx] = avw / (wd * coarse_area);
  }
  if (sd > 1.0e-6f) {
    c.us[idx] = aus / (sd * coarse_area);
    c.vs[idx] = avs / (sd * coarse_area);
  }
}

// Fine columns off the coarse window are parked (restrict_fine :8476-8485);
// their volume is booked as frozen, in FP64.
__global__ void k_park(FinePlanes f, FineAxes ax, double* frozen) {
  int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i >= kFineRes || j >= kFineRes) return;
  if (ax.col[i] >= 0 && ax.row[j] >= 0) return;
  size_t fidx = (size_t)j * kFineRes + i;
  float w = f.water[fidx], s = f.slurry[fidx];
  if (w != 0.0f) atomicAdd(&frozen[0], (double)fmaxf(w, 0.0f) * kFineArea);
  if (s != 0.0f) atomicAdd(&frozen[1], (double)fmaxf(s, 0.0f) * kFineArea);
}

__global__ void k_fill(float* p, float val, int n) {
  int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k < n) p[k] = val;
}

// ---- near field tiles (src/fluid_core/near.rs) ------------------------------

// World keys are stored with the sign bit of each half flipped, so the keys
// of the cells around the world origin ((-1, -1), (-1, -2)) do not land on
// the all-ones sentinels.  The two keys that still do (kx = 2^31 - 1,
// kz = 2^31 - 1 or 2^31 - 2, some 430,000 km out) keep no tile.
__device__ inline unsigned long long pack_key(int kx, int kz) {
  return ((unsigned long long)((unsigned)kx ^ 0x80000000u) << 32) | (unsigned long long)((unsigned)kz ^ 0x80000000u);
}

__device__ inline bool reserved_key(unsigned long long key) { return key == kEmptyKey || key == kTombKey; }

__device__ inline unsigned tile_hash(int kx, int kz, unsigned cap) {
  unsigned h = (unsigned)kx * 0x9E3779B1u ^ ((unsigned)kz * 0x85EBCA6Bu + 0x27d4eb2fu);
  h ^= h >> 15;
  return h & (cap - 1);
}

__device__ int tile_find(const NearTable& t, int kx, int kz) {
  unsigned long long key = pack_key(kx, kz);
  if (reserved_key(key)) return -1;
  unsigned s = tile_hash(kx, kz, t.cap);
  for (unsigned n = 0; n < t.cap; n++, s = (s + 1) & (t.cap - 1)) {
    unsigned long long k = t.keys[s];
    if (k == kEmptyKey) return -1;
    if (k == key) return (int)s;
  }
  return -1;
}

__device__ int tile_insert(NearTable& t, int kx, int kz) {
  int found = tile_find(t, kx, kz);
  if (found >= 0) return found;
  unsigned long long key = pack_key(kx, kz);
  if (reserved_key(key)) return -1;
  unsigned s = tile_hash(kx, kz, t.cap);
  for (unsigned n = 0; n < t.cap; n++, s = (s + 1) & (t.cap - 1)) {
    unsigned long long k = t.keys[s];
    if (k == kEmptyKey || k == kTombKey) {
      unsigned long long prev = atomicCAS(&t.keys[s], k, key);
      if (prev == k) {
        atomicAdd(t.count, 1u);
        return (int)s;
      }
    }
  }
  return -1;
}

__device__ void tile_remove(NearTable& t, int kx, int kz) {
  int s = tile_find(t, kx, kz);
  if (s < 0) return;
  t.keys[s] = kTombKey;
  atomicSub(t.count, 1u);
}

__device__ inline float tile_left(float idle) { return idle <= 0.0f ? 1.0f : expf(-idle / 1.5f); }

__device__ inline float abs_peak(const float* r) {
  float hi[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k = 0; k < 100; k += 4)
    for (int l = 0; l < 4; l++) {
      float a = fabsf(r[k + l]);
      hi[l] = a > hi[l] ? a : hi[l];
    }
  return fmaxf(fmaxf(fmaxf(hi[0], hi[1]), hi[2]), hi[3]);
}

__device__ inline float floor_scale(float base, const float* r) {
  float low = 0.0f;
  for (int k = 0; k < 100; k++) low = fminf(low, r[k]);
  if (low >= 0.0f || base + low >= 0.0f) return 1.0f;
  if (base <= 0.0f) return 0.0f;
  return clampf(base / -low, 0.0f, 1.0f);
}

__device__ inline void recentre(float* r) {
  double m = 0.0;
  for (int k = 0; k < 100; k++) m += (double)r[k];
  float mean = (float)(m / 100.0);
  for (int k = 0; k < 100; k++) r[k] = r[k] - mean;
}

// near_capture (near.rs:199-261), one held column per thread.
__global__ void k_near_capture(FinePlanes f, NearTable t, const NearColumn* cols, int ncols) {
  int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (q >= ncols) return;
  NearColumn nc = cols[q];
  float sw = 0.0f, ss = 0.0f;
  for (int b = 0; b < 10; b++) {
    size_t fr = (size_t)(nc.j0 + b) * kFineRes + nc.i0;
    float rw = -0.0f, rs = -0.0f;
    for (int a = 0; a < 10; a++) {
      rw = rw + f.water[fr + a];
      rs = rs + f.slurry[fr + a];
    }
    sw = sw + rw;
    ss = ss + rs;
  }
  if (sw <= 0.0f && ss <= 0.0f) {
    tile_remove(t, nc.kx, nc.kz);
    return;
  }
  float mw = sw / 100.0f, ms = ss / 100.0f;
  float rw[100], rs[100];
  for (int b = 0; b < 10; b++) {
    size_t fr = (size_t)(nc.j0 + b) * kFineRes + nc.i0;
    for (int a = 0; a < 10; a++) {
      rw[b * 10 + a] = f.water[fr + a] - mw;
      rs[b * 10 + a] = f.slurry[fr + a] - ms;
    }
  }
  float peak = fmaxf(abs_peak(rw), abs_peak(rs));
  if (peak <= 1.0e-5f && !nc.inner) {
    tile_remove(t, nc.kx, nc.kz);
    return;
  }
  recentre(rw);
  recentre(rs);
  int s = tile_insert(t, nc.kx, nc.kz);
  if (s < 0) return;
  float* d = t.data + (size_t)s * 400;
  for (int k = 0; k < 100; k++) {
    d[k] = rw[k];
    d[100 + k] = rs[k];
  }
  if (nc.inner && sw > 1.0e-10f) {
    float mu = 0.0f, mv = 0.0f;
    for (int b = 0; b < 10; b++) {
      size_t fr = (size_t)(nc.j0 + b) * kFineRes + nc.i0;
      for (int a = 0; a < 10; a++) {
        float w = f.water[fr + a];
        mu = mu + w * f.uw[fr + a];
        mv = mv + w * f.vw[fr + a];
      }
    }
    float uw = mu / sw, vw = mv / sw;
    for (int b = 0; b < 10; b++) {
      size_t fr = (size_t)(nc.j0 + b) * kFineRes + nc.i0;
      for (int a = 0; a < 10; a++) {
        bool wet = f.water[fr + a] > 1.0e-6f;
        d[200 + b * 10 + a] = wet ? f.uw[fr + a] - uw : 0.0f;
        d[300 + b * 10 + a] = wet ? f.vw[fr + a] - vw : 0.0f;
      }
    }
  } else {
    for (int k = 0; k < 100; k++) {
      d[200 + k] = 0.0f;
      d[300 + k] = 0.0f;
    }
  }
  t.peak[s] = peak;
  t.idle[s] = 0.0f;
  t.fresh[s] = 1;
}

// near_apply's lay-down (near.rs:265-296), one held column per thread.
__global__ void k_near_apply(FinePlanes f, NearTable t, const NearColumn* cols, int ncols) {
  int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (q >= ncols) return;
  NearColumn nc = cols[q];
  int s = tile_find(t, nc.kx, nc.kz);
  if (s < 0) return;
  const float* d = t.data + (size_t)s * 400;
  size_t f0 = (size_t)nc.j0 * kFineRes + nc.i0;
  float bw = f.water[f0], bs = f.slurry[f0];
  float left = tile_left(t.idle[s]);
  float aw = left * floor_scale(bw / fmaxf(left, 1.0e-6f), d);
  float as = left * floor_scale(bs / fmaxf(left, 1.0e-6f), d + 100);
  for (int b = 0; b < 10; b++)
    for (int a = 0; a < 10; a++) {
      size_t fi = (size_t)(nc.j0 + b) * kFineRes + nc.i0 + a;
      int k = b * 10 + a;
      float w = fmaxf(bw + aw * d[k], 0.0f);
      float sv = fmaxf(bs + as * d[100 + k], 0.0f);
      f.water[fi] = w;
      f.slurry[fi] = sv;
      if (nc.inner && w > 1.0e-6f) {
        f.uw[fi] = f.uw[fi] + left * d[200 + k];
        f.vw[fi] = f.vw[fi] + left * d[300 + k];
      }
    }
}

// near_capture with one 128-lane workgroup per held column: lane k < 100
// owns fine cell (k / 10, k % 10) of the column.  The sums the reference
// makes in order (row sums then their sum, the FP64 mean, the
// momentum-weighted velocity) run on lane 0 over LDS copies in the same
// order; the peaks are maxima (order-free once NaN is dropped, as the
// one-lane scan drops it).
__global__ __launch_bounds__(128) void k_near_capture_wg(FinePlanes f, NearTable t, const NearColumn* cols) {
  __shared__ float w_s[100], s_s[100], u_s[100], v_s[100];
  __shared__ float bc_f[4];
  __shared__ float pk[2][4];
  __shared__ int bc_i[2];
  const NearColumn nc = cols[blockIdx.x];
  const int k = (int)threadIdx.x;
  const bool cellk = k < 100;
  const size_t fi = cellk ? (size_t)(nc.j0 + k / 10) * kFineRes + nc.i0 + k % 10 : 0;
  float w = 0.0f, sl = 0.0f;
  if (cellk) {
    w = f.water[fi];
    sl = f.slurry[fi];
    w_s[k] = w;
    s_s[k] = sl;
  }
  __syncthreads();
  if (k == 0) {
    float sw = 0.0f, ss = 0.0f;
    for (int b = 0; b < 10; b++) {
      float rw = -0.0f, rs = -0.0f;
      for (int a = 0; a < 10; a++) {
        rw = rw + w_s[b * 10 + a];
        rs = rs + s_s[b * 10 + a];
      }
      sw = sw + rw;
      ss = ss + rs;
    }
    bc_f[0] = sw;
    bc_f[1] = ss;
    const bool empty = sw <= 0.0f && ss <= 0.0f;
    bc_i[0] = empty;
    if (empty) tile_remove(t, nc.kx, nc.kz);
  }
  __syncthreads();
  if (bc_i[0]) return;
  const float sw = bc_f[0], ss = bc_f[1];
  const float mw = sw / 100.0f, ms = ss / 100.0f;
  const float rw = w - mw, rs = sl - ms;
  // abs_peak: the one-lane scan skips NaN (a > hi is false), so drop it here
  // and take plain maxima.
  float aw = cellk ? fabsf(rw) : 0.0f, as = cellk ? fabsf(rs) : 0.0f;
  aw = aw > 0.0f ? aw : 0.0f;
  as = as > 0.0f ? as : 0.0f;
  for (int o = 32; o > 0; o >>= 1) {
    aw = fmaxf(aw, __shfl_xor(aw, o));
    as = fmaxf(as, __shfl_xor(as, o));
  }
  if ((k & 63) == 0) {
    pk[0][k >> 6] = aw;
    pk[1][k >> 6] = as;
  }
  if (cellk) {
    w_s[k] = rw;
    s_s[k] = rs;
  }
  __syncthreads();
  const float peak = fmaxf(fmaxf(pk[0][0], pk[0][1]), fmaxf(pk[1][0], pk[1][1]));
  if (peak <= 1.0e-5f && !nc.inner) {
    if (k == 0) tile_remove(t, nc.kx, nc.kz);
    return;
  }
  if (k == 0) {
    // recentre (near.rs) in the reference's order, FP64.
    double m = 0.0, n = 0.0;
    for (int q = 0; q < 100; q++) {
      m += (double)w_s[q];
      n += (double)s_s[q];
    }
    bc_f[2] = (float)(m / 100.0);
    bc_f[3] = (float)(n / 100.0);
    bc_i[1] = tile_insert(t, nc.kx, nc.kz);
  }
  __syncthreads();
  const int slot = bc_i[1];
  if (slot < 0) return;
  float* d = t.data + (size_t)slot * 400;
  if (cellk) {
    d[k] = rw - bc_f[2];
    d[100 + k] = rs - bc_f[3];
  }
  if (nc.inner && sw > 1.0e-10f) {
    float uu = 0.0f, vv = 0.0f;
    if (cellk) {
      uu = f.uw[fi];
      vv = f.vw[fi];
      u_s[k] = w * uu;
      v_s[k] = w * vv;
    }
    __syncthreads();
    if (k == 0) {
      float mu = 0.0f, mv = 0.0f;
      for (int q = 0; q < 100; q++) {
        mu = mu + u_s[q];
        mv = mv + v_s[q];
      }
      bc_f[0] = mu / sw;
      bc_f[1] = mv / sw;
    }
    __syncthreads();
    if (cellk) {
      const bool wet = w > 1.0e-6f;
      d[200 + k] = wet ? uu - bc_f[0] : 0.0f;
      d[300 + k] = wet ? vv - bc_f[1] : 0.0f;
    }
  } else if (cellk) {
    d[200 + k] = 0.0f;
    d[300 + k] = 0.0f;
  }
  if (k == 0) {
    t.peak[slot] = peak;
    t.idle[slot] = 0.0f;
    t.fresh[slot] = 1;
  }
}

// near_apply with one 128-lane workgroup per held column (lane k < 100 owns
// fine cell k); floor_scale's minima are order-free (fminf drops NaN).
__global__ __launch_bounds__(128) void k_near_apply_wg(FinePlanes f, NearTable t, const NearColumn* cols,
                                                      double* handoff) {
  __shared__ float lo[2][2];
  __shared__ int slot_s;
  const NearColumn nc = cols[blockIdx.x];
  const int k = (int)threadIdx.x;
  const bool cellk = k < 100;
  if (k == 0) slot_s = tile_find(t, nc.kx, nc.kz);
  __syncthreads();
  const int s = slot_s;
  if (s < 0) return;
  const float* d = t.data + (size_t)s * 400;
  float dw = cellk ? d[k] : 0.0f, ds = cellk ? d[100 + k] : 0.0f;
  float mw = fminf(0.0f, dw), ms = fminf(0.0f, ds);
  for (int o = 32; o > 0; o >>= 1) {
    mw = fminf(mw, __shfl_xor(mw, o));
    ms = fminf(ms, __shfl_xor(ms, o));
  }
  if ((k & 63) == 0) {
    lo[0][k >> 6] = mw;
    lo[1][k >> 6] = ms;
  }
  __syncthreads();
  const size_t f0 = (size_t)nc.j0 * kFineRes + nc.i0;
  const float bw = f.water[f0], bs = f.slurry[f0];
  const float left = tile_left(t.idle[s]);
  auto scale = [&](float base, float low) -> float {
    if (low >= 0.0f || base + low >= 0.0f) return 1.0f;
    if (base <= 0.0f) return 0.0f;
    return clampf(base / -low, 0.0f, 1.0f);
  };
  const float aw = left * scale(bw / fmaxf(left, 1.0e-6f), fminf(lo[0][0], lo[0][1]));
  const float as = left * scale(bs / fmaxf(left, 1.0e-6f), fminf(lo[1][0], lo[1][1]));
  // Every lane has read the column's first cell above; the writes below may
  // overwrite it.
  __syncthreads();
  double gain_w = 0.0, gain_s = 0.0;
  if (cellk) {
    const size_t fi = (size_t)(nc.j0 + k / 10) * kFineRes + nc.i0 + k % 10;
    const float was_w = f.water[fi], was_s = f.slurry[fi];
    const float w = fmaxf(bw + aw * dw, 0.0f);
    const float sv = fmaxf(bs + as * ds, 0.0f);
    f.water[fi] = w;
    f.slurry[fi] = sv;
    if (nc.inner && w > 1.0e-6f) {
      f.uw[fi] = f.uw[fi] + left * d[200 + k];
      f.vw[fi] = f.vw[fi] + left * d[300 + k];
    }
    gain_w = ((double)w - (double)fmaxf(was_w, 0.0f)) * kFineArea;
    gain_s = ((double)sv - (double)fmaxf(was_s, 0.0f)) * kFineArea;
  }
  if (!handoff) return;
  for (int o = 32; o > 0; o >>= 1) {
    gain_w += __shfl_xor(gain_w, o);
    gain_s += __shfl_xor(gain_s, o);
  }
  if ((k & 63) == 0) {
    if (gain_w != 0.0) atomicAdd(&handoff[0], gain_w);
    if (gain_s != 0.0) atomicAdd(&handoff[1], gain_s);
  }
}

// near_fade (near.rs:300-313), one table slot per thread.
__global__ void k_near_fade(NearTable t, float dt) {
  unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= t.cap) return;
  unsigned long long k = t.keys[s];
  if (k == kEmptyKey || k == kTombKey) return;
  if (t.fresh[s]) {
    t.fresh[s] = 0;
    return;
  }
  t.idle[s] = t.idle[s] + dt;
  if (!(t.peak[s] * tile_left(t.idle[s]) > 1.0e-5f)) {
    t.keys[s] = kTombKey;
    atomicSub(t.count, 1u);
  }
}

// Live tiles move into a fresh table (drops tombstones).
__global__ void k_near_rehash(NearTable from, NearTable to) {
  unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= from.cap) return;
  unsigned long long k = from.keys[s];
  if (k == kEmptyKey || k == kTombKey) return;
  int kx = (int)((unsigned)(k >> 32) ^ 0x80000000u), kz = (int)((unsigned)(k & 0xffffffffull) ^ 0x80000000u);
  int d = tile_insert(to, kx, kz);
  if (d < 0) return;
  for (int q = 0; q < 400; q++) to.data[(size_t)d * 400 + q] = from.data[(size_t)s * 400 + q];
  to.peak[d] = from.peak[s];
  to.idle[d] = from.idle[s];
  to.fresh[d] = from.fresh[s];
}

__global__ void k_near_clear(NearTable t) {
  unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= t.cap) return;
  t.keys[s] = kEmptyKey;
  t.fresh[s] = 0;
  if (s == 0) *t.count = 0u;
}

// coarse_cell + publish (src/fluid_core.rs:6998-7034,
// src/render_contract.rs:424-496): film (water, slurry), flow (u, v, speed,
// turbidity) and, when soil planes are given, ground_rut
// (bed + placed, rut, bed, placed).
__global__ void k_publish(CorePlanes c, const float* soil_bed, const float* soil_placed, const float* soil_rut,
                          const float* susp, const float* turbid, float* ground_rut4, float* film2, float* flow4,
                          int n) {
  int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= n) return;
  auto fin = [](float x) { return isfinite(x) ? x : 0.0f; };
  float water = fin(c.water[k]), slurry = fin(c.slurry[k]);
  float film = water + slurry;
  float u = 0.0f, v = 0.0f;
  if (!(film < 1.0e-5f)) {
    u = fin((c.uw[k] * water + c.us[k] * slurry) / film);
    v = fin((c.vw[k] * water + c.vs[k] * slurry) / film);
  }
  bool dry = !isfinite(water) || !isfinite(slurry) || water + slurry < 1.0e-5f;
  if (dry) u = v = 0.0f;
  float speed = dry ? 0.0f : fin(sqrtf(u * u + v * v));
  if (film2) {
    film2[2 * k] = water;
    film2[2 * k + 1] = slurry;
  }
  if (flow4) {
    flow4[4 * k] = u;
    flow4[4 * k + 1] = v;
    flow4[4 * k + 2] = speed;
    // coarse_cell's turbidity (src/fluid_core.rs:7025-7032): wheel-lifted
    // plus stirred solids over the water, clamped; the contract's channel
    // (src/render_contract.rs:476-481) clamps it again.
    float turb = 0.0f;
    if (!(water < 1.0e-5f)) {
      const float load = fmaxf(fin(susp ? susp[k] : 0.0f), 0.0f) + fmaxf(fin(turbid ? turbid[k] : 0.0f), 0.0f);
      float tb = load / water;
      tb = tb < 0.0f ? 0.0f : (tb > 1.0f ? 1.0f : tb);
      tb = fin(tb);
      turb = tb < 0.0f ? 0.0f : (tb > 1.0f ? 1.0f : tb);
    }
    flow4[4 * k + 3] = turb;
  }
  if (ground_rut4) {
    float bed = soil_bed ? soil_bed[k] : c.bed[k];
    float placed = soil_placed ? soil_placed[k] : 0.0f;
    float rut = soil_rut ? soil_rut[k] : 0.0f;
    ground_rut4[4 * k] = fin(bed + placed);
    ground_rut4[4 * k + 1] = fin(rut);
    ground_rut4[4 * k + 2] = fin(bed);
    ground_rut4[4 * k + 3] = fin(placed);
  }
}

// The render contract's moisture, wet_edge and foam_crest maps
// (src/render_contract.rs:457-500, :526-569): one lane per texel, reading
// its four neighbours' films and wet flags directly.
__device__ inline float fin0(float x) { return isfinite(x) ? x : 0.0f; }

__device__ inline float surface_moisture(const CorePlanes& c, const float* wetness, int k) {
  float m = fin0(c.moisture[k]);
  return wetness ? fmaxf(m, wetness[k]) : m;
}

__device__ inline bool surface_wet(const CorePlanes& c, const float* wetness, int k) {
  // is_wet (src/render_contract.rs:261-263) on coarse_cell's finite films.
  return surface_moisture(c, wetness, k) >= 0.45f || fin0(c.water[k]) + fin0(c.slurry[k]) >= 0.004f;
}

__global__ void k_publish_surface(CorePlanes c, const float* wetness, const float* placed, const float* foam,
                                  int res, float* moisture1, float* wet_edge1, float* foam2) {
  const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= res * res) return;
  const int i = k % res, j = k / res;
  if (moisture1) moisture1[k] = fin0(surface_moisture(c, wetness, k));
  const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
  if (wet_edge1) {
    const bool w = surface_wet(c, wetness, k);
    float edge = 0.0f;
    for (int q = 0; q < 4; q++) {
      const int ni = i + di[q], nj = j + dj[q];
      if (ni < 0 || nj < 0 || ni >= res || nj >= res) continue;
      if (w != surface_wet(c, wetness, nj * res + ni)) {
        edge = 1.0f;
        break;
      }
    }
    wet_edge1[k] = edge;
  }
  if (foam2) {
    const float h = fin0(c.water[k]) + fin0(c.slurry[k]);
    float sum = 0.0f, count = 0.0f;
    for (int q = 0; q < 4; q++) {
      const int ni = i + di[q], nj = j + dj[q];
      if (ni < 0 || nj < 0 || ni >= res || nj >= res) continue;
      const int n = nj * res + ni;
      sum += fin0(c.water[n]) + fin0(c.slurry[n]);
      count += 1.0f;
    }
    const float mean = count > 0.0f ? sum / count : h;
    const float excess = fin0(fmaxf(h - mean, 0.0f));
    const float pl = placed ? fin0(fmaxf(placed[k], 0.0f)) : 0.0f;
    // The foam tracer where the film is at least 4 mm (render_contract.rs:563-566).
    const float f = foam ? fin0(foam[k]) : 0.0f;
    foam2[2 * k] = h >= 0.004f ? fminf(fmaxf(f, 0.0f), 1.0f) : 0.0f;
    foam2[2 * k + 1] = fin0(pl + excess);
  }
}

// pull_substrate (src/world.rs:5511-5537): the sheet's bed is the pressed
// ground, bed + placed; the rut map is the cut below the initial ground.
__global__ void k_ground(const float* soil_bed, const float* soil_placed, const float* soil_init, float* sheet_bed,
                         float* rut, int n) {
  int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= n) return;
  sheet_bed[k] = soil_bed[k] + soil_placed[k];
  if (rut) rut[k] = fmaxf(soil_init[k] - soil_bed[k], 0.0f);
}

inline dim3 grid2(int w, int h, dim3 b) {
  return dim3((unsigned)((w + (int)b.x - 1) / (int)b.x), (unsigned)((h + (int)b.y - 1) / (int)b.y));
}

}  // namespace

void core_push_on(const GridGeo& g, float* water, float* slurry, float* uw, float* vw, float* us, float* vs,
                  const PushList& pl, int i0, int j0, int w, int h, hipStream_t st) {
  if (w <= 0 || h <= 0 || pl.n <= 0) return;
  dim3 b(32, 8);
  hipLaunchKernelGGL(k_push_on, grid2(w, h, b), b, 0, st, g, water, slurry, uw, vw, us, vs, pl, i0, j0, w, h);
}

void core_push_on_many(const GridGeo& g, float* water, float* slurry, float* uw, float* vw, float* us, float* vs,
                       const PushBox* boxes, const CorePush* pushes, int n, int max_w, int max_h, hipStream_t st) {
  if (n <= 0 || max_w <= 0 || max_h <= 0) return;
  dim3 b(32, 8);
  dim3 gr = grid2(max_w, max_h, b);
  gr.z = (unsigned)n;
  hipLaunchKernelGGL(k_push_on_many, gr, b, 0, st, g, water, slurry, uw, vw, us, vs, boxes, pushes);
}

void core_peel(const GridGeo& g, float* water, const float* uw, const float* vw, const PushList& pl, int i0, int j0,
               int w, int h, const uint8_t* hold, const GridGeo& cg, bool skip_held, SprayEvent* ev, unsigned* nev,
               unsigned cap, double* vol, double area, hipStream_t st) {
  if (w <= 0 || h <= 0 || pl.n <= 0) return;
  dim3 b(32, 8);
  hipLaunchKernelGGL(k_peel, grid2(w, h, b), b, 0, st, g, water, uw, vw, pl, i0, j0, w, h, hold, cg, (int)skip_held, ev,
                     nev, cap, vol, area);
}

void core_soften(float* u, float* v, int res, hipStream_t st) {
  dim3 b(32, 8);
  hipLaunchKernelGGL(k_soften, grid2(res, res, b), b, 0, st, u, v, res);
}

void core_nu(float* nu, const uint8_t* grain, const float* moisture, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_nu, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nu, grain, moisture, n);
}

__global__ void k_fill_bytes(uint8_t* p, uint8_t val, int n) {
  int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k < n) p[k] = val;
}

void core_fill_bytes(uint8_t* p, uint8_t val, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_bytes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, val, n);
}

void core_fill(float* p, float val, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, val, n);
}

void core_refine_copy(const CorePlanes& c, const FinePlanes& f, const uint8_t* hold, const FineAxes& ax, int ci0,
                      int cj0, int cw, int ch, hipStream_t st) {
  dim3 b(32, 8);
  hipLaunchKernelGGL(k_refine, grid2(kFineRes, kFineRes, b), b, 0, st, c, f, hold, ax);
  if (cw > 0 && ch > 0) hipLaunchKernelGGL(k_zero_held, grid2(cw, ch, b), b, 0, st, c, hold, ci0, cj0, cw, ch);
}

void core_restrict(const CorePlanes& c, const FinePlanes& f, const uint8_t* hold, const CoarseSpans& sp,
                   const FineAxes& ax, double* frozen, double* handoff, hipStream_t st) {
  dim3 b(16, 16);
  if (sp.ncol > 0 && sp.nrow > 0)
    hipLaunchKernelGGL(k_restrict, grid2(sp.ncol, sp.nrow, b), b, 0, st, c, f, hold, sp, handoff);
  hipLaunchKernelGGL(k_park, grid2(kFineRes, kFineRes, b), b, 0, st, f, ax, frozen);
}

void core_near_capture(const FinePlanes& f, const NearTable& t, const NearColumn* cols, int ncols, hipStream_t st) {
  if (ncols <= 0) return;
  hipLaunchKernelGGL(k_near_capture_wg, dim3((unsigned)ncols), dim3(128), 0, st, f, t, cols);
}

void core_near_apply(const FinePlanes& f, const NearTable& t, const NearColumn* cols, int ncols, double* handoff,
                     hipStream_t st) {
  if (ncols <= 0) return;
  hipLaunchKernelGGL(k_near_apply_wg, dim3((unsigned)ncols), dim3(128), 0, st, f, t, cols, handoff);
}

void core_near_fade(const NearTable& t, float dt, hipStream_t st) {
  hipLaunchKernelGGL(k_near_fade, dim3((t.cap + 255) / 256), dim3(256), 0, st, t, dt);
}

void core_near_clear(const NearTable& t, hipStream_t st) {
  hipLaunchKernelGGL(k_near_clear, dim3((t.cap + 255) / 256), dim3(256), 0, st, t);
}

void core_near_rehash(const NearTable& from, const NearTable& to, hipStream_t st) {
  core_near_clear(to, st);
  hipLaunchKernelGGL(k_near_rehash, dim3((from.cap + 255) / 256), dim3(256), 0, st, from, to);
}

void core_ground(const float* soil_bed, const float* soil_placed, const float* soil_init, float* sheet_bed, float* rut,
                 int n, hipStream_t st) {
  hipLaunchKernelGGL(k_ground, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, soil_bed, soil_placed, soil_init,
                     sheet_bed, rut, n);
}

void core_publish(const CorePlanes& c, const float* soil_bed, const float* soil_placed, const float* soil_rut,
                  const float* susp, const float* turbid, float* ground_rut4, float* film2, float* flow4,
                  hipStream_t st) {
  int n = kWindowRes * kWindowRes;
  hipLaunchKernelGGL(k_publish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, c, soil_bed, soil_placed, soil_rut,
                     susp, turbid, ground_rut4, film2, flow4, n);
}

// ---- walls: the mask and the scour ------------------------------------------

// solid_overlaps (src/fluid_core.rs:4584-4604) at a cell centre.
__device__ inline bool solid_overlaps_at(const SolidDesc& s, float x, float z, float eta, float bed) {
  const float dx = x - s.x, dz = z - s.z;
  const float right_x = s.fwd_z, right_z = -s.fwd_x;
  const float along = dx * s.fwd_x + dz * s.fwd_z;
  const float lat = dx * right_x + dz * right_z;
  if (!(fabsf(along) <= fmaxf(s.half_length, 0.0f) && fabsf(lat) <= fmaxf(s.half_width, 0.0f))) return false;
  return eta > s.bottom_y + 1.0e-4f && bed < s.top_y;
}

__global__ void k_solid_mask(const float* __restrict__ water, const float* __restrict__ bed, int res, float cell,
                             float ox, float oz, const SolidDesc* __restrict__ descs, int n,
                             uint8_t* __restrict__ mask) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i >= res || j >= res) return;
  const size_t idx = (size_t)j * res + i;
  const float x = ox + ((float)i + 0.5f) * cell;
  const float z = oz + ((float)j + 0.5f) * cell;
  // The free surface beside the cell: the highest of it and its 4-neighbours.
  float eta = bed[idx] + water[idx];
  if (i > 0) eta = fmaxf(eta, bed[idx - 1] + water[idx - 1]);
  if (i + 1 < res) eta = fmaxf(eta, bed[idx + 1] + water[idx + 1]);
  if (j > 0) eta = fmaxf(eta, bed[idx - res] + water[idx - res]);
  if (j + 1 < res) eta = fmaxf(eta, bed[idx + res] + water[idx + res]);
  uint8_t m = 0;
  for (int k = 0; k < n; k++) {
    const SolidDesc s = descs[k];
    if (i < s.i0 || i > s.i1 || j < s.j0 || j > s.j1) continue;
    if (solid_overlaps_at(s, x, z, eta, bed[idx])) m = (uint8_t)s.value;
  }
  mask[idx] = m;
}

// scour_contacts: a wet open cell touching a contact (value 1) whose
// shear 1000 * 0.03 * |u|^2 exceeds 1.5 Pa lifts bed into the slurry, at
// most 0.08 m in all.  Each cell reads and writes only itself (and the mask).
__global__ void k_scour(CorePlanes c, const uint8_t* __restrict__ mask, float* __restrict__ bed_drop, float dt,
                        double* eroded, unsigned* hole_bits, double* source) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), j = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  const int res = kWindowRes;
  double gain = 0.0;
  if (i >= 1 && j >= 1 && i < res - 1 && j < res - 1) {
    const size_t idx = (size_t)j * res + i;
    if (mask[idx] == 0 && !(c.water[idx] < 0.02f) &&
        (mask[idx - 1] == 1 || mask[idx + 1] == 1 || mask[idx - res] == 1 || mask[idx + res] == 1)) {
      const float speed = hypotf(c.uw[idx], c.vw[idx]);
      if (isfinite(speed)) {
        const float tau = 1000.0f * 0.03f * speed * speed;
        const float tau_c = 1.5f;
        if (tau > tau_c) {
          const float rate = fminf(0.0035f * (tau / tau_c - 1.0f), 0.012f);
          const float room = fmaxf(0.08f - bed_drop[idx], 0.0f);
          const float dh = fminf(rate * dt, room);
          if (dh > 0.0f) {
            const float was = c.slurry[idx];
            c.bed[idx] = c.bed[idx] - dh;
            c.slurry[idx] = was + dh;
            c.us[idx] = c.uw[idx];
            c.vs[idx] = c.vw[idx];
            c.grain[idx] = 1;  // Grain::Silt
            c.moisture[idx] = fmaxf(c.moisture[idx], 0.9f);
            const float drop = bed_drop[idx] + dh;
            bed_drop[idx] = drop;
            atomicAdd(eroded, (double)(dh * (kCoarseCell * kCoarseCell)));
            atomicMax(hole_bits, __float_as_uint(drop));
            gain = ((double)(was + dh) - (double)fmaxf(was, 0.0f)) * kCoarseArea;
          }
        }
      }
    }
  }
  if (gain != 0.0) atomicAdd(source, gain);
}

void core_solid_mask(const float* water, const float* bed, int res, float cell, float ox, float oz,
                     const SolidDesc* descs, int n, uint8_t* mask, hipStream_t st) {
  dim3 b(32, 8);
  hipLaunchKernelGGL(k_solid_mask, grid2(res, res, b), b, 0, st, water, bed, res, cell, ox, oz, descs, n, mask);
}

void core_scour(const CorePlanes& c, const uint8_t* mask, float* bed_drop, float dt, double* eroded,
                unsigned* hole_bits, double* source, hipStream_t st) {
  dim3 b(32, 8);
  hipLaunchKernelGGL(k_scour, grid2(kWindowRes, kWindowRes, b), b, 0, st, c, mask, bed_drop, dt, eroded, hole_bits,
                     source);
}

void core_publish_surface(const CorePlanes& c, const float* wetness, const float* placed, const float* foam, int res,
                          float* moisture1, float* wet_edge1, float* foam2, hipStream_t st) {
  const int n = res * res;
  hipLaunchKernelGGL(k_publish_surface, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, c, wetness, placed, foam,
                     res, moisture1, wet_edge1, foam2);
}

}  // namespace silt
