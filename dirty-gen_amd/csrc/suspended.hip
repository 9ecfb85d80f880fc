// Wheel-lifted suspended solids on the 20 cm window: entrain_wake and
// step_suspended / advect_suspended (src/fluid_core.rs:8750-9003).
//
// Both passes are one workgroup of 1024 lanes.  entrain_wake walks the
// pushes in order (a later tire lifts from the column the earlier one left)
// with the cells of one push's footprint on separate lanes.  The advection
// runs in phases: every cell lets its solids fall out; every donor in the
// plume's box finds its give and receiver; every receiver takes its donors'
// gives in the donors' raster order (the order the reference walks them,
// which decides how much room the later donor finds); every cell then
// subtracts what was accepted from it, adds what it took, and is capped at
// its water.  Per cell the arithmetic is the reference's f32 in its order,
// except the fall-out factor exp(-dt w_s / h), which is exp in double
// rounded to f32 (within one ulp of a libm expf).  The entrained and settled
// accounts, f32 running sums of the lifts and falls in the reference, book
// here the FP64 change of the plane each pass makes: the lift's, and the
// column's loss over a step (what fell out, what left the window, and the
// f32 rounding of the carry), so column + settled + parked = entrained holds
// to the FP64 sums.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "tracers.hpp"

namespace silt {

namespace {

constexpr int R = kWindowRes;
constexpr int RR = kWindowRes * kWindowRes;
constexpr int kT = 1024;
constexpr float kFall = 0.004f;  // SUSP_FALL_M_S, src/fluid_core.rs:122

__device__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int t = (int)threadIdx.x;
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int k = 0; k < kT / 64; k++) s += red[k];
  __syncthreads();
  return s;
}

__device__ inline float clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

__global__ __launch_bounds__(kT) void k_entrain(SuspPlanes sp, const float* __restrict__ water, TracerPushes tp,
                                                float dt, float ox, float oz) {
  __shared__ double red[kT / 64];
  const int t = (int)threadIdx.x;
  const float cell = kCoarseCell;
  double lifted = 0.0;
  for (int k = 0; k < tp.n; k++) {
    const CorePush p = tp.p[k];
    if (p.speed < 0.3f) continue;
    const float fl = fmaxf(sqrtf(p.fx * p.fx + p.fz * p.fz), 1.0e-4f);
    const float fx = p.fx / fl, fz = p.fz / fl;
    const float rx = fz, rz = -fx;
    const float reach = p.radius * 3.4f + cell;
    const int i0 = max((int)floorf(((p.x - reach) - ox) / cell), 0);
    const int i1 = min((int)floorf(((p.x + reach) - ox) / cell), R - 1);
    const int j0 = max((int)floorf(((p.z - reach) - oz) / cell), 0);
    const int j1 = min((int)floorf(((p.z + reach) - oz) / cell), R - 1);
    const float aft = p.radius * 3.2f;
    const float half = p.width * 0.55f + 0.08f;
    const int w = i1 - i0 + 1, h = j1 - j0 + 1;
    if (w > 0 && h > 0)
      for (int q = t; q < w * h; q += kT) {
        const int i = i0 + q % w, j = j0 + q / w;
        const float x = ox + ((float)i + 0.5f) * cell;
        const float z = oz + ((float)j + 0.5f) * cell;
        const float dx = x - p.x, dz = z - p.z;
        const float along = dx * fx + dz * fz;
        if (along > -0.05f || along < -aft) continue;
        const float lat = fabsf(dx * rx + dz * rz);
        if (lat > half) continue;
        const int idx = j * R + i;
        const float wv = water[idx];
        if (!isfinite(wv) || wv < 0.004f) continue;
        const float along_t = clamp01((-along) / aft);
        const float lat_t = clamp01(lat / half);
        const float core = (1.0f - along_t) * (1.0f - lat_t);
        const float room = fmaxf(wv - sp.susp[idx], 0.0f);
        const float add = fminf(4.0f * dt * core * wv, room);
        if (add <= 1.0e-8f) continue;
        const float before = sp.susp[idx], after = before + add;
        sp.susp[idx] = after;
        lifted += (double)after - (double)before;
      }
    __syncthreads();
  }
  lifted = block_sum(lifted, red);
  if (t == 0 && lifted != 0.0) atomicAdd(&sp.acc[0], lifted * kCoarseArea);
}

__global__ __launch_bounds__(kT) void k_susp_step(SuspPlanes sp, const float* __restrict__ water,
                                                  const float* __restrict__ vw, float dt) {
  __shared__ double red[kT / 64];
  __shared__ int box[4];
  const int t = (int)threadIdx.x;
  const float cell = kCoarseCell;
  float* susp = sp.susp;
  if (t == 0) {
    box[0] = R;
    box[1] = -1;
    box[2] = R;
    box[3] = -1;
  }
  __syncthreads();
  // solids fall out of the column (:8822-8863)
  double settled = 0.0;
  int bi0 = R, bi1 = -1, bj0 = R, bj1 = -1;
  for (int idx = t; idx < RR; idx += kT) {
    float s = susp[idx];
    if (!isfinite(s) || s <= 1.0e-12f) {
      if (isfinite(s)) settled += (double)s;
      susp[idx] = 0.0f;
      continue;
    }
    const float w = water[idx];
    if (!isfinite(w) || w < 1.0e-5f || s > w) {
      settled += (double)s;
      susp[idx] = 0.0f;
      continue;
    }
    const float keep = (float)exp((double)(-dt * kFall / w));
    const float landed = s * (1.0f - keep);
    if (landed > 0.0f) {
      const float left = s - landed;
      settled += (double)s - (double)left;
      s = left;
      susp[idx] = s;
    }
    if (s > 1.0e-8f) {
      const int i = idx % R, j = idx / R;
      bi0 = min(bi0, i);
      bi1 = max(bi1, i);
      bj0 = min(bj0, j);
      bj1 = max(bj1, j);
    }
  }
  if (bi1 >= 0) {
    atomicMin(&box[0], bi0);
    atomicMax(&box[1], bi1);
    atomicMin(&box[2], bj0);
    atomicMax(&box[3], bj1);
  }
  __syncthreads();
  const int i0 = max(box[0] - 1, 0), i1 = min(box[1] + 1, R - 1);
  const int j0 = max(box[2] - 1, 0), j1 = min(box[3] + 1, R - 1);
  const bool any = box[1] >= 0;
  const int bw = i1 - i0 + 1, bh = j1 - j0 + 1;
  if (any) {
    // every donor's give along v and where it goes (:8882-8916): dir +1 to
    // the row below (j + 1), -1 to the row above, 2 off the window, 0 none
    for (int k = t; k < bw * bh; k += kT) {
      const int i = i0 + k % bw, j = j0 + k / bw, idx = j * R + i;
      int8_t dir = 0;
      float give = 0.0f, conc = 0.0f;
      const float depth = susp[idx], wv = water[idx];
      if (!(depth <= 1.0e-12f || !isfinite(depth) || !isfinite(wv) || wv < 1.0e-5f)) {
        // film_uv (:3490-3499) of the water alone; u is 0 off a driven surface
        const float hh = wv + 0.0f;
        float v = 0.0f;
        if (isfinite(hh) && !(hh < 1.0e-5f)) {
          v = (vw[idx] * wv + 0.0f * 0.0f) / hh;
          v = isfinite(v) ? v : 0.0f;
        }
        conc = depth / wv;
        float fv = fabsf(v) * dt / cell;
        const float sum = 0.0f + fv;
        if (sum > 0.45f) fv *= 0.45f / sum;
        give = depth * fv;
        if (give > 1.0e-12f) {
          if (v > 0.0f)
            dir = j + 1 < R ? 1 : 2;
          else
            dir = j > 0 ? -1 : 2;
        }
      }
      sp.dir[idx] = dir;
      sp.give[idx] = give;
      sp.conc[idx] = conc;
    }
    __syncthreads();
    // every receiver takes its donors' gives in their raster order (:8917-8935):
    // the donor above (giving down) before the donor below (giving up)
    const int rj0 = max(j0 - 1, 0), rj1 = min(j1 + 1, R - 1);
    const int rh = rj1 - rj0 + 1;
    for (int k = t; k < bw * rh; k += kT) {
      const int i = i0 + k % bw, j = rj0 + k / bw, n = j * R + i;
      const float wr = water[n];
      const bool ok_r = isfinite(wr) && !(wr < 1.0e-5f);
      float tk = 0.0f;
      for (int side = 0; side < 2; side++) {
        const int d = side == 0 ? n - R : n + R;
        const int jd = side == 0 ? j - 1 : j + 1;
        if (jd < j0 || jd > j1 || sp.dir[d] != (side == 0 ? 1 : -1)) continue;
        float a = 0.0f;
        if (ok_r) {
          const float have = susp[n] + tk;
          const float room = fmaxf(sp.conc[d] * wr - have, 0.0f);
          const float got = fminf(fminf(sp.give[d], room), fmaxf(wr - have, 0.0f));
          if (!(got <= 1.0e-12f)) {
            a = got;
            tk += got;
          }
        }
        sp.accept[d] = a;
      }
      sp.taken[n] = tk;
    }
    __syncthreads();
    // what left each cell, what it took, and the cap at its water (:8936-8999)
    for (int k = t; k < bw * bh; k += kT) {
      const int i = i0 + k % bw, j = j0 + k / bw, idx = j * R + i;
      const float s0 = susp[idx];
      float s = s0;
      const int8_t dir = sp.dir[idx];
      if (dir == 2) {
        s -= sp.give[idx];
      } else if (dir != 0 && sp.accept[idx] > 0.0f) {
        s -= sp.accept[idx];
      }
      s += sp.taken[idx];
      if (!isfinite(s) || s <= 0.0f) {
        s = 0.0f;
      } else {
        const float w = water[idx];
        if (!isfinite(w) || w < 1.0e-5f)
          s = 0.0f;
        else if (s > w)
          s = w;
      }
      settled += (double)s0 - (double)s;
      susp[idx] = s;
    }
  }
  settled = block_sum(settled, red);
  if (t == 0 && settled != 0.0) atomicAdd(&sp.acc[1], settled * kCoarseArea);
}

}  // namespace

void susp_entrain(const SuspPlanes& sp, const float* water, const TracerPushes& tp, float dt, float ox, float oz,
                  hipStream_t st) {
  hipLaunchKernelGGL(k_entrain, dim3(1), dim3(kT), 0, st, sp, water, tp, dt, ox, oz);
}

void susp_step(const SuspPlanes& sp, const float* water, const float* vw, float dt, hipStream_t st) {
  hipLaunchKernelGGL(k_susp_step, dim3(1), dim3(kT), 0, st, sp, water, vw, dt);
}

}  // namespace silt
