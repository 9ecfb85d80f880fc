// Work boxes of a sheet found on the GPU: sheet_work_boxes
// (src/fluid_core.rs:1325-1447) for a wet set that is not one rectangle.
//
// The reference labels 4-connected wet bodies by row runs with a union-find
// on the host.  Here the same idea runs on the GPU over the wet bounding box:
//
//   1. one wavefront per row finds the row's runs (ballot over 64 columns at a
//      time) and points every wet cell at its run's first cell (its head);
//   2. each pair of overlapping runs of neighbouring rows is joined once (at
//      the first column of their overlap, where one of the two cells starts
//      its run) by a lock-free union of heads: links always point to a
//      smaller index and only ever decrease (atomicMin), so finds need no
//      locks and the forest has no cycles; finds halve their path as they go.
//      Rows join inside blocks of 32 first, the heads are compressed, then
//      the block boundaries join: a mostly wet sheet would otherwise chain
//      one head per row and every find would walk the chain;
//   3. heads are compressed to their root, then every cell reads its head's
//      root; each root takes a body number and every cell widens its body's
//      box with atomics.
//
// Links join heads only, so a find walks at most one link per row of a body.
// Only the body boxes (with the body's first cell, to restore the
// reference's row-scan order) come back to the host, which grows them by 6,
// merges touching ones and applies the reference's one-or-more-than-8 rule.
#include <limits.h>
#include <string.h>

#include <algorithm>
#include <utility>
#include <vector>

#include "plan.hpp"
#include "sheet_fused.hpp"

namespace silt {

namespace {

constexpr int kBodyStride = 5;  // i0, i1, j0, j1, first cell

__device__ inline bool wet_at(const float* h, int c, float eps) { return h[c] > eps; }

// Root of x with path halving.  Every store only lowers a link (atomicMin),
// so a concurrent union that lowered it first is never undone.
__device__ inline int uf_find(int* parent, int x) {
  for (;;) {
    int p = parent[x];
    if (p == x) return x;
    int g = parent[p];
    if (g != p) atomicMin(&parent[x], g);
    x = p;
  }
}

__device__ inline void uf_unite(int* parent, int a, int b) {
  for (;;) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return;
    if (a > b) {
      int t = a;
      a = b;
      b = t;
    }
    // Link root b under the smaller a.  If b stopped being a root meanwhile,
    // its new parent still belongs with a: continue from there.
    int old = atomicMin(&parent[b], a);
    if (old == b) return;
    b = old;
  }
}

// Step 1: one 64-lane wavefront per row of the box.  parent[c] = the head of
// c's run for wet cells, -1 for dry ones.  Cells outside the wet box are
// never read (the box holds every wet cell).
__global__ __launch_bounds__(256) void k_label_runs(const float* __restrict__ h, int* __restrict__ parent, int res,
                                                    Box bb, float eps) {
  const int lane = (int)(threadIdx.x & 63);
  const int j = bb.j0 + (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (j > bb.j1) return;  // whole wavefronts leave together
  const int row = j * res;
  int carry = -1;  // head of the run that reaches the previous segment's end
  for (int s = bb.i0; s <= bb.i1; s += 64) {
    const int i = s + lane;
    const bool in = i <= bb.i1;
    const bool w = in && wet_at(h, row + i, eps);
    const unsigned long long wm = __ballot(w);
    // A run starts at a wet cell whose left neighbour in this segment is dry
    // (bit 0: always a start here; it continues the carried run instead when
    // the previous segment ended wet).
    const unsigned long long starts = wm & ~(wm << 1);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long mine = starts & upto;
    int head = -1;  // stays -1 only on dry lanes
    if (mine) {
      const int b = 63 - __clzll((long long)mine);
      head = (b == 0 && carry >= 0) ? carry : row + s + b;
    }
    if (in) parent[row + i] = w ? head : -1;
    const int last = __shfl(head, 63);
    carry = ((wm >> 63) & 1ull) ? last : -1;
  }
}

// Step 2: join overlapping runs of rows j-1 and j, in two phases so no find
// walks a long chain: phase 0 joins rows inside blocks of kJoinBlock rows
// (chains stay within a block), the heads are compressed, phase 1 joins the
// block boundaries (chains of block roots).
constexpr int kJoinBlock = 32;

__global__ void k_label_join(const float* __restrict__ h, int* parent, int res, Box bb, float eps, int phase) {
  const int i = bb.i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = bb.j0 + 1 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i > bb.i1 || j > bb.j1) return;
  if ((((j - bb.j0) % kJoinBlock) == 0) != (phase == 1)) return;
  const int c = j * res + i;
  if (!wet_at(h, c, eps) || !wet_at(h, c - res, eps)) return;
  const bool left_in = i > bb.i0;
  const bool starts_here = !left_in || !wet_at(h, c - 1, eps);
  const bool starts_above = !left_in || !wet_at(h, c - res - 1, eps);
  if (!starts_here && !starts_above) return;
  uf_unite(parent, parent[c], parent[c - res]);
}

// Step 3a: heads point straight at their root.
__global__ void k_label_heads(int* parent, int res, Box bb) {
  const int i = bb.i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = bb.j0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i > bb.i1 || j > bb.j1) return;
  const int c = j * res + i;
  if (parent[c] < 0) return;
  if (i == bb.i0 || parent[c - 1] < 0) parent[c] = uf_find(parent, c);
}

// Step 3b: every cell takes its head's root (heads are final here, and a
// head's slot already holds that root).
__global__ void k_label_cells(int* parent, int res, Box bb) {
  const int i = bb.i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = bb.j0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i > bb.i1 || j > bb.j1) return;
  const int c = j * res + i;
  const int p = parent[c];
  if (p < 0 || p == c) return;
  parent[c] = parent[p];
}

// Step 3c: number the roots.  A root's own slot becomes -(id + 2).
__global__ void k_label_roots(int* parent, int* bodies, unsigned* count, int cap, int res, Box bb) {
  const int i = bb.i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = bb.j0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i > bb.i1 || j > bb.j1) return;
  const int c = j * res + i;
  if (parent[c] != c) return;
  const unsigned id = atomicAdd(count, 1u);
  if (id < (unsigned)cap) {
    int* b = bodies + kBodyStride * id;
    b[0] = INT_MAX;
    b[1] = -1;
    b[2] = INT_MAX;
    b[3] = -1;
    b[4] = c;  // the body's first cell in row-major order (links point down)
  }
  parent[c] = -(int)id - 2;
}

__global__ void k_label_boxes(const int* parent, int* bodies, int cap, int res, Box bb) {
  const int i = bb.i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = bb.j0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i > bb.i1 || j > bb.j1) return;
  const int c = j * res + i;
  const int r = parent[c];
  if (r == -1) return;
  // Only the ends of a run can widen its body's box (every row of a body
  // has a run, so the first cells give the rows too): a few atomics per run
  // instead of four per cell on one hot address.
  const bool first = i == bb.i0 || parent[c - 1] == -1;
  const bool last = i == bb.i1 || parent[c + 1] == -1;
  if (!first && !last) return;
  const int code = r < -1 ? r : parent[r];
  const int id = -code - 2;
  if (id >= cap) return;
  int* b = bodies + kBodyStride * id;
  if (first) {
    atomicMin(&b[0], i);
    atomicMin(&b[2], j);
    atomicMax(&b[3], j);
  }
  if (last) atomicMax(&b[1], i);
}

__global__ void k_zero_u32(unsigned* p, int n) {
  int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k < n) p[k] = 0u;
}

dim3 grid_over(int w, int hgt, dim3 blk) {
  return dim3((unsigned)((w + (int)blk.x - 1) / (int)blk.x), (unsigned)((hgt + (int)blk.y - 1) / (int)blk.y));
}

Box halo3_of(const SheetStats& st, int res) {
  return Box{std::max(st.i0 - 3, 0), std::min(st.i1 + 3, res - 1), std::max(st.j0 - 3, 0),
             std::min(st.j1 + 3, res - 1)};
}

}  // namespace

int plan_begin(const float* h, int res, float cell, float dt, float eps, const SheetStats& st, PlanScratch& ps,
               StepPlan* plan, hipStream_t stream) {
  *plan = StepPlan{};
  ps.labelled = false;
  if (!(dt > 0.0f) || st.wet == 0) return 1;
  if (cell <= kFineCell + 1.0e-6f) return 0;  // 2 cm sheets plan on awake tiles
  const long long area = (long long)(st.i1 - st.i0 + 1) * (long long)(st.j1 - st.j0 + 1);
  if ((long long)st.wet == area) {
    plan->boxes.push_back(halo3_of(st, res));
    return 1;
  }
  const size_t n = (size_t)res * res;
  if (ps.cap_cells < n) {
    if (ps.parent) (void)hipFree(ps.parent);
    ps.parent = nullptr;
    ps.cap_cells = 0;
    if (hipMalloc(&ps.parent, n * sizeof(int)) != hipSuccess) return SILT_ENOMEM;
    ps.cap_cells = n;
  }
  // Layout: [count, 3 pad][bodies...]; one copy brings the count and the
  // first kPlanBodyPrefix bodies.
  const size_t body_bytes = 16 + (size_t)kPlanBodyCap * kBodyStride * sizeof(int);
  const size_t head_bytes = 16 + (size_t)kPlanBodyPrefix * kBodyStride * sizeof(int);
  if (!ps.bodies) {
    if (hipMalloc(&ps.bodies, body_bytes) != hipSuccess ||
        hipHostMalloc(&ps.h_bodies, body_bytes, hipHostMallocDefault) != hipSuccess)
      return SILT_ENOMEM;
  }
  unsigned* d_count = reinterpret_cast<unsigned*>(ps.bodies);
  int* d_list = ps.bodies + 4;
  const Box bb{st.i0, st.i1, st.j0, st.j1};
  const int bw = bb.i1 - bb.i0 + 1, bh = bb.j1 - bb.j0 + 1;
  const dim3 blk(64, 4);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, d_count, 1);
  hipLaunchKernelGGL(k_label_runs, dim3((unsigned)((bh + 3) / 4)), dim3(256), 0, stream, h, ps.parent, res, bb, eps);
  if (bh > 1) {
    hipLaunchKernelGGL(k_label_join, grid_over(bw, bh - 1, blk), blk, 0, stream, h, ps.parent, res, bb, eps, 0);
    if (bh > kJoinBlock) {
      hipLaunchKernelGGL(k_label_heads, grid_over(bw, bh, blk), blk, 0, stream, ps.parent, res, bb);
      hipLaunchKernelGGL(k_label_join, grid_over(bw, bh - 1, blk), blk, 0, stream, h, ps.parent, res, bb, eps, 1);
    }
  }
  hipLaunchKernelGGL(k_label_heads, grid_over(bw, bh, blk), blk, 0, stream, ps.parent, res, bb);
  hipLaunchKernelGGL(k_label_cells, grid_over(bw, bh, blk), blk, 0, stream, ps.parent, res, bb);
  hipLaunchKernelGGL(k_label_roots, grid_over(bw, bh, blk), blk, 0, stream, ps.parent, d_list, d_count,
                     kPlanBodyCap, res, bb);
  hipLaunchKernelGGL(k_label_boxes, grid_over(bw, bh, blk), blk, 0, stream, ps.parent, d_list, kPlanBodyCap, res,
                     bb);
  if (hipMemcpyAsync(ps.h_bodies, ps.bodies, head_bytes, hipMemcpyDeviceToHost, stream) != hipSuccess)
    return SILT_EDEVICE;
  ps.labelled = true;
  ps.halo3 = halo3_of(st, res);
  return 1;
}

int plan_boxes(PlanScratch& ps, int res, StepPlan* plan, hipStream_t stream) {
  if (!ps.labelled) return 1;
  ps.labelled = false;
  const unsigned count = reinterpret_cast<const unsigned*>(ps.h_bodies)[0];
  if (count > (unsigned)kPlanBodyCap) return 0;  // too many bodies to read back: plan on the host
  if (count > (unsigned)kPlanBodyPrefix &&
      (hipMemcpyAsync(ps.h_bodies + 4, ps.bodies + 4, (size_t)count * kBodyStride * sizeof(int),
                      hipMemcpyDeviceToHost, stream) != hipSuccess ||
       hipStreamSynchronize(stream) != hipSuccess))
    return SILT_EDEVICE;
  // Body numbers come from atomics; the reference lists bodies in the
  // row-major order of their first cell, so sort on that.
  std::vector<std::pair<int, Box>> found((size_t)count);
  for (unsigned k = 0; k < count; k++) {
    const int* b = ps.h_bodies + 4 + kBodyStride * k;
    found[k] = {b[4], Box{b[0], b[1], b[2], b[3]}};
  }
  std::sort(found.begin(), found.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
  std::vector<Box> bodies;
  bodies.reserve(found.size());
  for (auto& f : found) bodies.push_back(f.second);
  plan->boxes = boxes_from_bodies(std::move(bodies), res, ps.halo3);
  return 1;
}

int plan_nsub_begin(const float* u, const float* v, int res, float cell, float dt, const SheetStats& st,
                    PlanScratch& ps, StepPlan* plan, hipStream_t stream) {
  ps.maxs2_pending = false;
  if (plan->boxes.empty()) return 0;
  float s2_all;
  memcpy(&s2_all, &st.max_s2_bits, 4);
  plan->nsub = nsub_from_s2(s2_all, dt, cell);
  // The largest speed anywhere bounds the boxes'; a bound of one substep is
  // exact, otherwise the boxes' own maximum decides.
  if (plan->nsub == 1) return 0;
  if (!ps.maxs2) {
    if (hipMalloc(&ps.maxs2, sizeof(unsigned)) != hipSuccess ||
        hipHostMalloc(&ps.h_maxs2, sizeof(unsigned), hipHostMallocDefault) != hipSuccess)
      return SILT_ENOMEM;
  }
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, ps.maxs2, 1);
  for (const Box& b : plan->boxes) box_max_s2(u, v, res, b, ps.maxs2, stream);
  if (hipMemcpyAsync(ps.h_maxs2, ps.maxs2, sizeof(unsigned), hipMemcpyDeviceToHost, stream) != hipSuccess)
    return SILT_EDEVICE;
  ps.maxs2_pending = true;
  return 0;
}

void plan_nsub_end(PlanScratch& ps, float cell, float dt, StepPlan* plan) {
  if (!ps.maxs2_pending) return;
  ps.maxs2_pending = false;
  float s2;
  memcpy(&s2, ps.h_maxs2, 4);
  plan->nsub = nsub_from_s2(s2, dt, cell);
}

int plan_sheet_device(const float* h, const float* u, const float* v, int res, float cell, float dt, float eps,
                      const SheetStats& st, PlanScratch& ps, StepPlan* plan, hipStream_t stream) {
  int rc = plan_begin(h, res, cell, dt, eps, st, ps, plan, stream);
  if (rc <= 0) return rc;
  if (ps.labelled) {
    if (hipStreamSynchronize(stream) != hipSuccess) return SILT_EDEVICE;
    if ((rc = plan_boxes(ps, res, plan, stream)) <= 0) return rc;
  }
  if ((rc = plan_nsub_begin(u, v, res, cell, dt, st, ps, plan, stream)) < 0) return rc;
  if (ps.maxs2_pending && hipStreamSynchronize(stream) != hipSuccess) return SILT_EDEVICE;
  plan_nsub_end(ps, cell, dt, plan);
  return 1;
}

void plan_scratch_free(PlanScratch& ps) {
  if (ps.parent) (void)hipFree(ps.parent);
  if (ps.bodies) (void)hipFree(ps.bodies);
  if (ps.h_bodies) (void)hipHostFree(ps.h_bodies);
  if (ps.maxs2) (void)hipFree(ps.maxs2);
  if (ps.h_maxs2) (void)hipHostFree(ps.h_maxs2);
  ps = PlanScratch{};
}

}  // namespace silt
