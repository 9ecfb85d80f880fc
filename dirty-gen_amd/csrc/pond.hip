// Ponds in tire cuts on the FluidCore window: pond_ruts (src/fluid_core.rs:
// 5509-5611), seat_rut_water (:5622-5736) and shove_rut_slurry (:5740-5792).
//
// Each pass is one workgroup of 1024 lanes over the 200 x 200 window, in
// phases separated by barriers.  Per-cell work runs on every lane; what the
// reference computes in an order that decides the f32 result runs in that
// order:
//  * the summed-area planes (prefix_plane, :1502-1516): one lane per row runs
//    the row sum, then one lane per column adds the rows top to bottom;
//  * the f32 totals (needs, rooms, dropped water, floor depth): each tile of
//    1024 cells compacts its non-zero terms in cell order into LDS and lane 0
//    adds them to the running total (a zero term would leave it unchanged);
//  * the slurry shoved off the cuts: each cut's ring and share are set first,
//    then every receiving cell adds the shares of the cuts around it in the
//    cuts' raster order, which is the order the reference walks its list.
// The pond's free surface median is a radix select over the floor cells.
#include "pond.hpp"

namespace silt {

namespace {

constexpr int kN = kWindowRes;
constexpr int kNN = kWindowRes * kWindowRes;
constexpr int kStride = kWindowRes + 1;
constexpr int kT = 1024;
constexpr int kWaves = kT / 64;
constexpr int kShoveRing = 4;  // SHOVE_RING, src/fluid_core.rs:481

struct Lds {
  float list[kT];
  int cnt[kWaves];
  double red[kWaves];
  unsigned hist[256];
  float total;
  int flag;
  unsigned key, mask;
  int kth;
  int box[4];
};

__device__ inline float area() { return kCoarseCell * kCoarseCell; }

// Adds the non-zero terms of one tile (one per lane, in lane order) to acc on
// lane 0, in cell order.  Every lane must call it.
__device__ float ordered_tile(Lds& l, float term, bool nz, float acc) {
  const int t = (int)threadIdx.x, w = t >> 6, lane = t & 63;
  const uint64_t m = __ballot(nz);
  if (lane == 0) l.cnt[w] = __popcll(m);
  __syncthreads();
  int base = 0;
  for (int k = 0; k < w; k++) base += l.cnt[k];
  if (nz) l.list[base + __popcll(m & ((1ull << lane) - 1ull))] = term;
  __syncthreads();
  if (t == 0) {
    int n = 0;
    for (int k = 0; k < kWaves; k++) n += l.cnt[k];
    for (int k = 0; k < n; k++) acc += l.list[k];
  }
  __syncthreads();
  return acc;
}

// FP64 sum over the workgroup, returned on every lane.
__device__ double block_sum(Lds& l, double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int t = (int)threadIdx.x;
  if ((t & 63) == 0) l.red[t >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < kWaves; k++) s += l.red[k];
  __syncthreads();
  return s;
}

__device__ inline void span(int i, int j, int rad, int* i0, int* i1, int* j0, int* j1) {
  *j0 = j - rad < 0 ? 0 : j - rad;
  *j1 = j + rad > kN - 1 ? kN - 1 : j + rad;
  *i0 = i - rad < 0 ? 0 : i - rad;
  *i1 = i + rad > kN - 1 ? kN - 1 : i + rad;
}

// plane_window (:1518-1525).
__device__ inline float plane_window(const float* p, int i0, int i1, int j0, int j1) {
  float a = p[j0 * kStride + i0], b = p[j0 * kStride + (i1 + 1)];
  float c = p[(j1 + 1) * kStride + i0], d = p[(j1 + 1) * kStride + (i1 + 1)];
  return d - b - c + a;
}

// Order-preserving key of a float (ascending keys = ascending values).
__device__ inline unsigned fkey(float x) {
  unsigned u = __float_as_uint(x);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float fkey_value(unsigned k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

__global__ __launch_bounds__(kT) void k_pond_ruts(CorePlanes c, PondScratch s, PondResult* r, double* handoff) {
  __shared__ Lds l;
  const int t = (int)threadIdx.x;
  const float A = area();
  float* water = c.water;
  const float* bed = c.bed;
  // any_rut (:5515-5524)
  bool any = false;
  for (int idx = t; idx < kNN; idx += kT) any = any || (isfinite(bed[idx]) && bed[idx] < -0.012f);
  if (!__syncthreads_or(any)) {
    if (t == 0) r->moved = 0.0f;
    return;
  }
  // still cells and their free surface (:5528-5537)
  uint8_t* still = s.m0;
  float* eta_of = s.f4;
  for (int idx = t; idx < kNN; idx += kT) {
    float b = bed[idx], w = water[idx];
    bool st = isfinite(b) && isfinite(w) && b > -0.008f && b < 0.03f && w > 0.018f;
    still[idx] = st ? 1 : 0;
    eta_of[idx] = st ? b + w : 0.0f;
    s.old_w[idx] = w;
  }
  __syncthreads();
  // prefix_plane of the mark and of the free surface: row sums ...
  if (t < kN) {
    float run = 0.0f;
    for (int i = 0; i < kN; i++) {
      run += still[t * kN + i] ? 1.0f : 0.0f;
      s.run[t * kN + i] = run;
    }
  } else if (t >= 256 && t < 256 + kN) {
    const int j = t - 256;
    float run = 0.0f;
    for (int i = 0; i < kN; i++) {
      run += eta_of[j * kN + i];
      s.run[kNN + j * kN + i] = run;
    }
  }
  __syncthreads();
  // ... then each column adds the rows above (the border stays zero).
  if (t < 2 * kStride) {
    const int plane = t / kStride, i = t % kStride;
    float* P = s.pre + plane * kStride * kStride;
    P[i] = 0.0f;
    if (i == 0) {
      for (int j = 1; j <= kN; j++) P[j * kStride] = 0.0f;
    } else {
      const float* R = s.run + plane * kNN;
      float acc = 0.0f;
      for (int j = 0; j < kN; j++) {
        acc = R[j * kN + (i - 1)] + acc;
        P[(j + 1) * kStride + i] = acc;
      }
    }
  }
  __syncthreads();
  // needs of the cuts (:5554-5578), their total in cell order
  const float* count_p = s.pre;
  const float* eta_p = s.pre + kStride * kStride;
  float total_need = 0.0f;
  for (int base = 0; base < kNN; base += kT) {
    const int idx = base + t;
    float extra = 0.0f;
    if (idx < kNN) {
      float b = bed[idx], w = water[idx];
      if (isfinite(b) && isfinite(w) && b < -0.012f) {
        int i0, i1, j0, j1;
        span(idx % kN, idx / kN, 10, &i0, &i1, &j0, &j1);
        float count = plane_window(count_p, i0, i1, j0, j1);
        float sum = plane_window(eta_p, i0, i1, j0, j1);
        if (!(count < 1.0f)) {
          float e = fmaxf((sum / count) - b - w, 0.0f);
          if (e > 1.0e-4f) extra = e;
        }
      }
      s.f0[idx] = extra;
    }
    total_need = ordered_tile(l, extra * A, extra > 0.0f, total_need);
  }
  if (t == 0) l.total = total_need;
  __syncthreads();
  total_need = l.total;
  if (!(total_need > 1.0e-8f)) {
    if (t == 0) r->moved = 0.0f;
    return;
  }
  // donors (:5582-5596)
  float total_room = 0.0f;
  for (int base = 0; base < kNN; base += kT) {
    const int idx = base + t;
    float room = 0.0f;
    if (idx < kNN) {
      if (still[idx]) {
        float rm = fmaxf(water[idx] - 0.012f, 0.0f);
        if (rm > 1.0e-4f) room = rm;
      }
      s.f1[idx] = room;
    }
    total_room = ordered_tile(l, room * A, room > 0.0f, total_room);
  }
  if (t == 0) l.total = total_room;
  __syncthreads();
  total_room = l.total;
  if (!(total_room > 1.0e-8f)) {
    if (t == 0) r->moved = 0.0f;
    return;
  }
  // move (:5597-5610); a cut is never a donor (bed < -0.012 vs > -0.008)
  const float scale = fminf(total_room / total_need, 1.0f);
  const float moved = total_need * scale;
  double dv = 0.0;
  for (int idx = t; idx < kNN; idx += kT) {
    float extra = s.f0[idx], room = s.f1[idx];
    float w0 = water[idx], w = w0;
    if (extra > 0.0f) w = w0 + extra * scale;
    if (room > 0.0f) {
      float give = moved * (room * A) / total_room;
      w = fmaxf(w0 - give / A, 0.0f);
    }
    if (w != w0) {
      water[idx] = w;
      dv += (double)w - (double)w0;
    }
  }
  dv = block_sum(l, dv);
  if (t == 0) {
    r->moved = moved;
    atomicAdd(&handoff[0], dv * kCoarseArea);
  }
}

__global__ __launch_bounds__(kT) void k_seat(CorePlanes c, PondScratch s, int gate, PondResult* r,
                                             double* handoff) {
  __shared__ Lds l;
  const int t = (int)threadIdx.x;
  const float A = area();
  float* water = c.water;
  float* slurry = c.slurry;
  float* pressed = s.f0;
  const uint8_t* rut = s.m1;
  uint8_t* floor_m = s.m0;
  // classify (:5634-5655); the gate of seat_rut_sheet (src/world.rs:5544-5546)
  int nfloor = 0;
  bool wet = false, any_rut = false;
  for (int idx = t; idx < kNN; idx += kT) {
    float g = pressed[idx];
    g = isfinite(g) ? g : 0.0f;
    pressed[idx] = g;
    const bool cut = rut[idx] != 0;
    const float w = water[idx];
    const bool on = !cut && g >= -0.02f && g < 0.04f && (w > 0.008f || c.moisture[idx] > 0.55f);
    floor_m[idx] = on ? 1 : 0;
    nfloor += on ? 1 : 0;
    wet = wet || !(w < 1.0e-5f) || !(slurry[idx] < 1.0e-5f);
    any_rut = any_rut || cut;
    s.old_w[idx] = w;
    s.old_s[idx] = slurry[idx];
    s.ring[idx] = 0;
    s.f1[idx] = 0.0f;
    s.f2[idx] = 0.0f;
    s.f3[idx] = 0.0f;
  }
  for (int off = 32; off > 0; off >>= 1) nfloor += __shfl_xor(nfloor, off);
  if (t == 0) {
    l.kth = 0;
    l.box[0] = kN;
    l.box[1] = -1;
    l.box[2] = kN;
    l.box[3] = -1;
  }
  __syncthreads();
  if ((t & 63) == 0) atomicAdd(&l.kth, nfloor);
  const int wet_any = __syncthreads_or(wet);
  const int rut_any = __syncthreads_or(any_rut);
  nfloor = l.kth;
  if (gate && !wet_any && !rut_any) {
    if (t == 0) r->has_floor = -1;
    return;
  }
  if (nfloor < 8) {
    if (t == 0) r->has_floor = 0;
    return;
  }
  // the median (select_nth_unstable_by at len / 2): radix select of the
  // floor cells' free surface, eight bits a pass
  unsigned prefix = 0, pmask = 0;
  int kth = nfloor / 2;
  for (int shift = 24; shift >= 0; shift -= 8) {
    if (t < 256) l.hist[t] = 0;
    __syncthreads();
    for (int idx = t; idx < kNN; idx += kT) {
      if (!floor_m[idx]) continue;
      unsigned k = fkey(pressed[idx] + water[idx]);
      if ((k & pmask) == prefix) atomicAdd(&l.hist[(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (t == 0) {
      int cum = 0, sel = 255;
      for (int b = 0; b < 256; b++) {
        if (cum + (int)l.hist[b] > kth) {
          sel = b;
          break;
        }
        cum += (int)l.hist[b];
      }
      l.kth = kth - cum;
      l.key = prefix | ((unsigned)sel << shift);
    }
    __syncthreads();
    kth = l.kth;
    prefix = l.key;
    pmask |= 255u << shift;
    __syncthreads();
  }
  const float eta = fkey_value(prefix);
  // cut_wet (:5660)
  bool cw = false;
  for (int idx = t; idx < kNN; idx += kT) cw = cw || (rut[idx] && water[idx] >= 0.012f);
  const int cut_wet = __syncthreads_or(cw);
  // cuts that open onto the floor within 8 cells: the column hangs from the
  // pond surface (:5674-5704)
  if (cut_wet) {
    int bi0 = kN, bi1 = -1, bj0 = kN, bj1 = -1;
    for (int idx = t; idx < kNN; idx += kT) {
      if (!rut[idx] || water[idx] < 0.012f) continue;
      const int i = idx % kN, j = idx / kN;
      int i0, i1, j0, j1;
      span(i, j, 8, &i0, &i1, &j0, &j1);
      int n = 0;
      for (int jj = j0; jj <= j1; jj++)
        for (int ii = i0; ii <= i1; ii++) n += floor_m[jj * kN + ii];
      if (n <= 0) continue;
      float w = water[idx];
      const float p = pressed[idx];
      const float limit = fmaxf(eta - p, 0.0f);
      if (w > limit + 0.004f) {
        s.f1[idx] = (w - limit) * A;
        w = limit;
        water[idx] = w;
      }
      const float room = fmaxf(eta - p - w, 0.0f);
      const float sl = slurry[idx];
      if (sl > room + 0.004f) {
        s.f3[idx] = sl - room;
        slurry[idx] = room;
        bi0 = min(bi0, i);
        bi1 = max(bi1, i);
        bj0 = min(bj0, j);
        bj1 = max(bj1, j);
      }
    }
    if (bi1 >= 0) {
      atomicMin(&l.box[0], bi0);
      atomicMax(&l.box[1], bi1);
      atomicMin(&l.box[2], bj0);
      atomicMax(&l.box[3], bj1);
    }
  }
  // water and slurry pushed up the bank (:5708-5722)
  for (int idx = t; idx < kNN; idx += kT) {
    const float p = pressed[idx];
    if (rut[idx] || p < 0.02f) continue;
    float w = water[idx];
    const float limit = fmaxf(eta - p, 0.0f);
    if (w > limit + 0.004f) {
      s.f2[idx] = (w - limit) * A;
      w = limit;
      water[idx] = w;
    }
    const float room = fmaxf(eta - p - w, 0.0f);
    if (slurry[idx] > room + 0.004f) slurry[idx] = room;
  }
  __syncthreads();
  // excess in the reference's order: the cuts' drops, then the bank's
  float excess = 0.0f;
  for (int base = 0; base < kNN; base += kT) {
    const int idx = base + t;
    const float v = idx < kNN ? s.f1[idx] : 0.0f;
    excess = ordered_tile(l, v, v > 0.0f, excess);
  }
  for (int base = 0; base < kNN; base += kT) {
    const int idx = base + t;
    const float v = idx < kNN ? s.f2[idx] : 0.0f;
    excess = ordered_tile(l, v, v > 0.0f, excess);
  }
  if (t == 0) l.total = excess;
  // shove_rut_slurry: each cut's nearest ring with a non-cut cell and its share
  const int bi0 = l.box[0], bi1 = l.box[1], bj0 = l.box[2], bj1 = l.box[3];
  __syncthreads();
  excess = l.total;
  if (bi1 >= 0) {
    const int bw = bi1 - bi0 + 1, bh = bj1 - bj0 + 1;
    for (int k = t; k < bw * bh; k += kT) {
      const int i = bi0 + k % bw, j = bj0 + k / bw, idx = j * kN + i;
      const float metres = s.f3[idx];
      if (!(metres > 0.0f)) continue;
      int ring = 0;
      unsigned count = 0;
      for (int d = 1; d <= kShoveRing && ring == 0; d++) {
        count = 0;
        const int ja = max(j - d, 0), jb = min(j + d, kN - 1), ia = max(i - d, 0), ib = min(i + d, kN - 1);
        for (int jj = ja; jj <= jb; jj++)
          for (int ii = ia; ii <= ib; ii++) {
            const int di = abs(ii - i), dj = abs(jj - j);
            if (max(di, dj) != d || rut[jj * kN + ii]) continue;
            count++;
          }
        if (count > 0) ring = d;
      }
      s.ring[idx] = (int8_t)ring;
      if (ring > 0)
        s.f4[idx] = metres / (float)count;
      else
        slurry[idx] += metres;
    }
    __syncthreads();
    // every receiving cell takes the shares of the cuts around it, in the
    // cuts' raster order
    const int ri0 = max(bi0 - kShoveRing, 0), ri1 = min(bi1 + kShoveRing, kN - 1);
    const int rj0 = max(bj0 - kShoveRing, 0), rj1 = min(bj1 + kShoveRing, kN - 1);
    const int rw = ri1 - ri0 + 1, rh = rj1 - rj0 + 1;
    for (int k = t; k < rw * rh; k += kT) {
      const int i = ri0 + k % rw, j = rj0 + k / rw, idx = j * kN + i;
      if (rut[idx]) continue;
      float sl = slurry[idx];
      uint8_t gr = c.grain[idx];
      bool touched = false;
      const int ja = max(j - kShoveRing, bj0), jb = min(j + kShoveRing, bj1);
      const int ia = max(i - kShoveRing, bi0), ib = min(i + kShoveRing, bi1);
      for (int jj = ja; jj <= jb; jj++)
        for (int ii = ia; ii <= ib; ii++) {
          const int e = jj * kN + ii;
          const int rg = s.ring[e];
          if (rg <= 0 || max(abs(ii - i), abs(jj - j)) != rg) continue;
          if (sl <= 1.0e-6f) gr = c.grain[e];
          sl += s.f4[e];
          touched = true;
        }
      if (touched) {
        slurry[idx] = sl;
        c.grain[idx] = gr;
      }
    }
  }
  __syncthreads();
  // the excess goes back onto the floor in proportion to depth (:5724-5734)
  if (excess > 1.0e-8f) {
    float room = 0.0f;
    for (int base = 0; base < kNN; base += kT) {
      const int idx = base + t;
      const float v = idx < kNN && floor_m[idx] ? fmaxf(water[idx], 0.0f) : 0.0f;
      room = ordered_tile(l, v, v != 0.0f, room);
    }
    if (t == 0) l.total = room;
    __syncthreads();
    room = l.total;
    if (room > 1.0e-6f)
      for (int idx = t; idx < kNN; idx += kT)
        if (floor_m[idx] && water[idx] > 0.0f) water[idx] += excess / A * (water[idx] / room);
  }
  double dw = 0.0, ds = 0.0;
  for (int idx = t; idx < kNN; idx += kT) {
    dw += (double)water[idx] - (double)s.old_w[idx];
    ds += (double)slurry[idx] - (double)s.old_s[idx];
  }
  dw = block_sum(l, dw);
  ds = block_sum(l, ds);
  if (t == 0) {
    r->eta = eta;
    r->has_floor = 1;
    atomicAdd(&handoff[0], dw * kCoarseArea);
    atomicAdd(&handoff[1], ds * kCoarseArea);
  }
}

__global__ void k_seat_classify(const float* __restrict__ bed, const float* __restrict__ placed,
                                const float* __restrict__ init, float* __restrict__ pressed, uint8_t* __restrict__ rut) {
  const int idx = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (idx >= kNN) return;
  const float b = bed[idx];
  pressed[idx] = b + (placed ? placed[idx] : 0.0f);
  rut[idx] = fmaxf(init[idx] - b, 0.0f) > 0.012f ? 1 : 0;
}

}  // namespace

void core_pond_ruts(const CorePlanes& c, const PondScratch& s, PondResult* r, double* handoff, hipStream_t st) {
  hipLaunchKernelGGL(k_pond_ruts, dim3(1), dim3(kT), 0, st, c, s, r, handoff);
}

void core_seat_rut_water(const CorePlanes& c, const PondScratch& s, int gate, PondResult* r, double* handoff,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_seat, dim3(1), dim3(kT), 0, st, c, s, gate, r, handoff);
}

void core_seat_classify(const float* soil_bed, const float* soil_placed, const float* soil_init, const PondScratch& s,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_seat_classify, dim3((kNN + 255) / 256), dim3(256), 0, st, soil_bed, soil_placed, soil_init,
                     s.f0, s.m1);
}

}  // namespace silt
