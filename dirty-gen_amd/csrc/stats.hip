// Plan statistics and ledger sums of a sheet (wet box, wet count, largest
// speed, FP64 volume): the reductions the step plan and ledger() need
// (sheet_work_boxes :1325-1347, the nsub scan :1722-1738, column_sum
// :4617-4620), done on the GPU with wavefront shuffles and one atomic per
// wave.
#include "sheet_fused.hpp"

#include <string.h>

#include <algorithm>

namespace silt {

namespace {

__global__ void k_stats_reset(SheetStats* s, int n) {
  int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= n) return;
  s[k].vol = 0.0;
  s[k].wet = 0ull;
  s[k].i0 = 0x7fffffff;
  s[k].i1 = -1;
  s[k].j0 = 0x7fffffff;
  s[k].j1 = -1;
  s[k].max_s2_bits = 0u;
  s[k].pad = 0;
}

// Grid-stride over rows: each block takes rows, each lane columns.
__global__ __launch_bounds__(256) void k_stats(const float* __restrict__ h, const float* __restrict__ u,
                                               const float* __restrict__ v, int res, float eps, SheetStats* s) {
  double vol = 0.0;
  unsigned wet = 0;
  int i0 = 0x7fffffff, i1 = -1, j0 = 0x7fffffff, j1 = -1;
  float s2 = 0.0f;
  for (int j = (int)blockIdx.x; j < res; j += (int)gridDim.x) {
    const float* hr = h + (size_t)j * res;
    const float* ur = u + (size_t)j * res;
    const float* vr = v + (size_t)j * res;
    // Four columns a blockDim apart per iteration: twelve loads in flight
    // per lane instead of three.
    const int step = (int)blockDim.x;
    for (int i = (int)threadIdx.x; i < res; i += 4 * step) {
      float hv[4], uu[4], vv[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int c = i + k * step;
        const bool in = c < res;
        hv[k] = in ? hr[c] : 0.0f;
        uu[k] = in ? ur[c] : 0.0f;
        vv[k] = in ? vr[c] : 0.0f;
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int c = i + k * step;
        vol += (double)fmaxf(hv[k], 0.0f);
        if (c < res && hv[k] > eps) {
          wet++;
          i0 = min(i0, c);
          i1 = max(i1, c);
          j0 = min(j0, j);
          j1 = max(j1, j);
        }
        s2 = fmaxf(s2, uu[k] * uu[k] + vv[k] * vv[k]);
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    vol += __shfl_xor(vol, off);
    wet += __shfl_xor(wet, off);
    i0 = min(i0, __shfl_xor(i0, off));
    i1 = max(i1, __shfl_xor(i1, off));
    j0 = min(j0, __shfl_xor(j0, off));
    j1 = max(j1, __shfl_xor(j1, off));
    s2 = fmaxf(s2, __shfl_xor(s2, off));
  }
  // The four wavefronts meet in LDS; one set of atomics per block (all
  // blocks' atomics on the same seven words otherwise queue up behind each
  // other).
  __shared__ double s_vol[4];
  __shared__ unsigned s_wet[4];
  __shared__ int s_box[4][4];
  __shared__ float s_s2[4];
  const int w = (int)(threadIdx.x >> 6);
  if ((threadIdx.x & 63) == 0) {
    s_vol[w] = vol;
    s_wet[w] = wet;
    s_box[w][0] = i0;
    s_box[w][1] = i1;
    s_box[w][2] = j0;
    s_box[w][3] = j1;
    s_s2[w] = s2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (int)(blockDim.x >> 6);
    for (int k = 1; k < nw; k++) {
      vol += s_vol[k];
      wet += s_wet[k];
      i0 = min(i0, s_box[k][0]);
      i1 = max(i1, s_box[k][1]);
      j0 = min(j0, s_box[k][2]);
      j1 = max(j1, s_box[k][3]);
      s2 = fmaxf(s2, s_s2[k]);
    }
    atomicAdd(&s->vol, vol);
    atomicAdd(&s->wet, (unsigned long long)wet);
    if (wet) {
      atomicMin(&s->i0, i0);
      atomicMax(&s->i1, i1);
      atomicMin(&s->j0, j0);
      atomicMax(&s->j1, j1);
    }
    atomicMax(&s->max_s2_bits, __float_as_uint(s2));
  }
}

__global__ __launch_bounds__(256) void k_box_max_s2(const float* __restrict__ u, const float* __restrict__ v, int res,
                                                    Box b, unsigned* out) {
  float s2 = 0.0f;
  for (int j = b.j0 + (int)blockIdx.x; j <= b.j1; j += (int)gridDim.x)
    for (int i = b.i0 + (int)threadIdx.x; i <= b.i1; i += (int)blockDim.x) {
      size_t idx = (size_t)j * res + i;
      float uu = u[idx], vv = v[idx];
      // Same operation order as the reference's scan: uu * uu + vv * vv.
      float q = uu * uu + vv * vv;
      s2 = q > s2 ? q : s2;
    }
  for (int off = 32; off > 0; off >>= 1) s2 = fmaxf(s2, __shfl_xor(s2, off));
  __shared__ float s_s2[4];
  if ((threadIdx.x & 63) == 0) s_s2[threadIdx.x >> 6] = s2;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) s2 = fmaxf(s2, s_s2[k]);
    atomicMax(out, __float_as_uint(s2));
  }
}

}  // namespace

void stats_reset(SheetStats* s, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_stats_reset, dim3(1), dim3(64), 0, st, s, n);
}

void stats_compute(const float* h, const float* u, const float* v, int res, float eps, SheetStats* s,
                   hipStream_t st) {
  stats_reset(s, 1, st);
  int blocks = res < 1024 ? res : 1024;
  hipLaunchKernelGGL(k_stats, dim3((unsigned)blocks), dim3(256), 0, st, h, u, v, res, eps, s);
}

void box_max_s2(const float* u, const float* v, int res, const Box& b, unsigned* out, hipStream_t st) {
  int rows = b.j1 - b.j0 + 1;
  int blocks = rows < 1024 ? rows : 1024;
  hipLaunchKernelGGL(k_box_max_s2, dim3((unsigned)blocks), dim3(256), 0, st, u, v, res, b, out);
}

}  // namespace silt

namespace silt {

// Substep count from the largest squared speed over the boxes
// (src/fluid_core.rs:1689-1760).
int nsub_from_s2(float s2, float dt, float cell) {
  float max_s = fmaxf(0.15f, sqrtf(s2));
  float q = ceilf((dt * max_s) / (0.42f * fmaxf(cell, 1.0e-3f)));
  return (int)clampf(q, 1.0f, 8.0f);
}

}  // namespace silt
