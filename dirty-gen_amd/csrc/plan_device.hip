// Work boxes of sheets found on the GPU: sheet_work_boxes
// (src/fluid_core.rs:1325-1447) for a wet set that is not one rectangle.
//
// The reference labels 4-connected wet bodies by row runs with a union-find
// on the host, grows every body's box by 6 cells, merges boxes that come
// within one cell of each other (boxes_touch, :1449-1456) and keeps the
// merged list unless it has one box or more than 8 (then the wet bounding
// box).  Only the merged boxes matter, and they are coarser than the bodies:
// two wet cells at most 14 cells apart on both axes always end in one merged
// box (their grown boxes touch).  So the GPU labels 4 x 4 TILES instead of
// cells:
//
//   * a tile is wet when it holds a wet cell; it keeps the tight box of its
//     wet cells and its first wet cell in row-major order;
//   * 4-connected wet tiles form clusters.  A body's cells sit in one cluster
//     (neighbouring cells are in the same or 4-neighbouring tiles), and a
//     cluster's cells sit in one merged box (cells of neighbouring tiles are
//     at most 7 apart), so a cluster is a union of bodies inside one merged
//     box;
//   * each cluster's tight box (of its wet cells) and first cell come back;
//     the host grows and merges them as the reference does its bodies.
//
// Merging is order-free in what it produces: every merge the reference makes
// from bodies also happens from clusters (boxes only grow), and nothing else
// does, so both end in the same set of boxes.  The list order can differ
// when 2 to 8 boxes remain; those boxes are at least 2 cells apart and the
// sheet steps them independently.
//
// The tile labelling is the run-based union-find the reference uses, on the
// GPU: one wavefront per tile row finds the runs (ballot over 64 tiles at a
// time); overlapping runs of neighbouring rows are joined once by a
// lock-free union of run heads (links only point to smaller indices and only
// decrease, so no locks and no cycles; finds halve their path), first inside
// blocks of 32 rows and then across the block boundaries, so a mostly wet
// sheet does not chain one head per row.  Every kernel serves a batch of
// sheets (a pair's two layers, every window of a batch): a block finds its
// sheet from the per-sheet block starts in its arguments.
#include <limits.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "plan.hpp"
#include "sheet_fused.hpp"

namespace silt {

namespace {

constexpr int kTile = 4;        // tile side in cells (<= 7 keeps clusters inside merged boxes)
constexpr int kBodyStride = 5;  // i0, i1, j0, j1, first cell
constexpr int kJoinBlock = 32;  // tile rows joined in the first phase

// Summary of a wet tile: bit 31 set, then the tight box of its wet cells in
// tile coordinates (li0, li1, lj0, lj1, 2 bits each) and the first wet
// column of row lj0 (2 bits).  0 for a dry tile.
constexpr unsigned kWetBit = 1u << 31;

__device__ inline unsigned tile_field(unsigned s, int k) { return (s >> (2 * k)) & 3u; }

// One sheet of a labelling batch (device copy of the host's table).
struct PlanJob {
  const float* h;    // cell depths, res x res
  int* parent;       // tile labels, tw x th
  unsigned* info;    // tile summaries, tw x th
  int res, tw, th;   // cells per side; tile grid over the wet box
  int ti0, tj0;      // first tile (grid coordinates / kTile)
  int bx;            // workgroups across the tile grid (64 x 4 tiles each)
  Box cb;            // wet cell box (statistics)
  float eps;
  int slot;          // body slot of this sheet
};

// Block starts of one launch: job k of the launch owns blocks
// [start[k], start[k + 1]).
constexpr int kLaunchJobs = 64;
struct JobStarts {
  int first;  // job index of start[0]
  int n;
  int start[kLaunchJobs + 1];
};

__device__ inline int job_of(const JobStarts& js, int b) {
  int lo = 0, hi = js.n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (js.start[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Body record (id) of sheet slot s: ids below kPlanBodyPrefix sit in the
// head region that comes back in one copy, the rest in the tail region.
__host__ __device__ inline size_t body_offset(int nslots, int s, int id) {
  const size_t counts = (size_t)((nslots + 63) & ~63);
  if (id < kPlanBodyPrefix) return counts + ((size_t)s * kPlanBodyPrefix + id) * kBodyStride;
  const size_t head = (size_t)nslots * kPlanBodyPrefix * kBodyStride;
  return counts + head + ((size_t)s * (kPlanBodyCap - kPlanBodyPrefix) + (id - kPlanBodyPrefix)) * kBodyStride;
}

// Tile coordinates of thread (threadIdx.x, threadIdx.y) of a 64 x 4 block
// of a job, or false when outside the job's tile grid.
__device__ inline bool tile_of(const JobStarts& js, const PlanJob* jobs, const PlanJob*& J, int& x, int& y) {
  const int b = (int)blockIdx.x;
  const int k = job_of(js, b);
  J = jobs + js.first + k;
  const int l = b - js.start[k];
  x = (l % J->bx) * 64 + (int)threadIdx.x;
  y = (l / J->bx) * 4 + (int)threadIdx.y;
  return x < J->tw && y < J->th;
}

// Tile summaries (and the sheet's body count cleared).
__global__ __launch_bounds__(256) void k_tile_summary(JobStarts js, const PlanJob* __restrict__ jobs,
                                                      int* __restrict__ bodies) {
  const PlanJob* J;
  int x, y;
  const bool in = tile_of(js, jobs, J, x, y);
  if (blockIdx.x == (unsigned)js.start[job_of(js, (int)blockIdx.x)] && threadIdx.x == 0 && threadIdx.y == 0)
    bodies[J->slot] = 0;
  if (!in) return;
  const int i0 = (J->ti0 + x) * kTile, j0 = (J->tj0 + y) * kTile;
  unsigned rows[kTile];
  unsigned cols = 0;
  int lj0 = -1, lj1 = -1;
#pragma unroll
  for (int r = 0; r < kTile; r++) {
    const int j = j0 + r;
    unsigned m = 0;
    if (j >= J->cb.j0 && j <= J->cb.j1) {
      const float* row = J->h + (size_t)j * J->res;
#pragma unroll
      for (int c = 0; c < kTile; c++) {
        const int i = i0 + c;
        if (i >= J->cb.i0 && i <= J->cb.i1 && row[i] > J->eps) m |= 1u << c;
      }
    }
    rows[r] = m;
    cols |= m;
    if (m) {
      if (lj0 < 0) lj0 = r;
      lj1 = r;
    }
  }
  unsigned s = 0;
  if (cols) {
    const int li0 = __ffs((int)cols) - 1, li1 = 31 - __clz((int)cols);
    const int fc = __ffs((int)rows[lj0]) - 1;
    s = kWetBit | (unsigned)li0 | ((unsigned)li1 << 2) | ((unsigned)lj0 << 4) | ((unsigned)lj1 << 6) |
        ((unsigned)fc << 8);
  }
  J->info[(size_t)y * J->tw + x] = s;
}

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

// Runs of wet tiles: one 64-lane wavefront per tile row.  parent[t] = the
// head of t's run for wet tiles, -1 for dry ones.
__global__ __launch_bounds__(256) void k_tile_runs(JobStarts js, const PlanJob* __restrict__ jobs) {
  const int b = (int)blockIdx.x;
  const int k = job_of(js, b);
  const PlanJob* J = jobs + js.first + k;
  const int lane = (int)(threadIdx.x & 63);
  const int y = (b - js.start[k]) * 4 + (int)(threadIdx.x >> 6);
  if (y >= J->th) return;  // whole wavefronts leave together
  const int row = y * J->tw;
  int carry = -1;  // head of the run that reaches the previous segment's end
  for (int s = 0; s < J->tw; s += 64) {
    const int x = s + lane;
    const bool in = x < J->tw;
    const bool w = in && J->info[row + x] != 0u;
    const unsigned long long wm = __ballot(w);
    // A run starts at a wet tile whose left neighbour in this segment is dry
    // (bit 0: continues the carried run when the previous segment ended wet).
    const unsigned long long starts = wm & ~(wm << 1);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const unsigned long long mine = starts & upto;
    int head = -1;
    if (mine) {
      const int bit = 63 - __clzll((long long)mine);
      head = (bit == 0 && carry >= 0) ? carry : row + s + bit;
    }
    if (in) J->parent[row + x] = w ? head : -1;
    const int last = __shfl(head, 63);
    carry = ((wm >> 63) & 1ull) ? last : -1;
  }
}

// Join overlapping runs of tile rows y-1 and y, once per pair (at the first
// column of the overlap, where one of the two tiles starts its run).  Phase
// 0 joins rows inside blocks of kJoinBlock rows, phase 1 the block
// boundaries after the heads were compressed.
__global__ __launch_bounds__(256) void k_tile_join(JobStarts js, const PlanJob* __restrict__ jobs, int phase) {
  const PlanJob* J;
  int x, y;
  if (!tile_of(js, jobs, J, x, y) || y == 0) return;
  if (((y % kJoinBlock) == 0) != (phase == 1)) return;
  const int t = y * J->tw + x, tw = J->tw;
  const unsigned* info = J->info;
  if (info[t] == 0u || info[t - tw] == 0u) return;
  const bool starts_here = x == 0 || info[t - 1] == 0u;
  const bool starts_above = x == 0 || info[t - tw - 1] == 0u;
  if (!starts_here && !starts_above) return;
  uf_unite(J->parent, J->parent[t], J->parent[t - tw]);
}

// Run heads point straight at their root.
__global__ __launch_bounds__(256) void k_tile_heads(JobStarts js, const PlanJob* __restrict__ jobs) {
  const PlanJob* J;
  int x, y;
  if (!tile_of(js, jobs, J, x, y)) return;
  const int t = y * J->tw + x;
  int* parent = J->parent;
  if (parent[t] < 0) return;
  if (x == 0 || parent[t - 1] < 0) parent[t] = uf_find(parent, t);
}

// Every tile takes its head's root (heads are final here).
__global__ __launch_bounds__(256) void k_tile_cells(JobStarts js, const PlanJob* __restrict__ jobs) {
  const PlanJob* J;
  int x, y;
  if (!tile_of(js, jobs, J, x, y)) return;
  const int t = y * J->tw + x;
  int* parent = J->parent;
  const int p = parent[t];
  if (p < 0 || p == t) return;
  parent[t] = parent[p];
}

// Number the roots (a root's slot becomes -(id + 2)) and open their records.
__global__ __launch_bounds__(256) void k_tile_roots(JobStarts js, const PlanJob* __restrict__ jobs, int* bodies,
                                                    int nslots) {
  const PlanJob* J;
  int x, y;
  if (!tile_of(js, jobs, J, x, y)) return;
  const int t = y * J->tw + x;
  if (J->parent[t] != t) return;
  const unsigned id = atomicAdd(reinterpret_cast<unsigned*>(bodies) + J->slot, 1u);
  if (id < (unsigned)kPlanBodyCap) {
    int* b = bodies + body_offset(nslots, J->slot, (int)id);
    b[0] = INT_MAX;
    b[1] = -1;
    b[2] = INT_MAX;
    b[3] = -1;
    b[4] = INT_MAX;
  }
  J->parent[t] = -(int)id - 2;
}

// Lowest (hi = false) or highest value of v over the lanes of the wavefront
// whose act is set, per cluster id: one atomic per distinct id in the
// wavefront (a cluster usually covers the whole wavefront).
__device__ inline void wave_extreme(int* bodies, int nslots, int slot, bool act, int id, int field, int v, bool hi) {
  unsigned long long todo = __ballot(act);
  const int lane = (int)(threadIdx.x & 63);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int lid = __shfl(id, leader);
    const bool mine = act && id == lid;
    int r = mine ? v : (hi ? INT_MIN : INT_MAX);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const int q = __shfl_xor(r, o);
      r = hi ? max(r, q) : min(r, q);
    }
    if (lane == leader) {
      int* b = bodies + body_offset(nslots, slot, lid);
      if (hi)
        atomicMax(b + field, r);
      else
        atomicMin(b + field, r);
    }
    todo &= ~__ballot(mine);
  }
}

// Cluster boxes in cells.  Only edge tiles can hold a cluster's extreme
// cells: the leftmost column of a cluster has dry tiles to its left (a wet
// left neighbour would be in the cluster), and so on; the first cell lies
// in the cluster's top tile row, whose tiles all have a dry tile above.
__global__ __launch_bounds__(256) void k_tile_boxes(JobStarts js, const PlanJob* __restrict__ jobs, int* bodies,
                                                    int nslots) {
  const PlanJob* J;
  int x, y;
  const bool in = tile_of(js, jobs, J, x, y);
  const int tw = J->tw;
  const int t = y * tw + x;
  const int r = in ? J->parent[t] : -1;
  int id = -1;
  if (r != -1) {
    const int code = r < -1 ? r : J->parent[r];
    id = -code - 2;
  }
  const bool ok = id >= 0 && id < kPlanBodyCap;
  bool left = false, right = false, top = false, bottom = false;
  int ci0 = 0, ci1 = 0, cj0 = 0, cj1 = 0, first = 0;
  if (ok) {
    const unsigned s = J->info[t];
    left = x == 0 || J->parent[t - 1] == -1;
    right = x == tw - 1 || J->parent[t + 1] == -1;
    top = y == 0 || J->parent[t - tw] == -1;
    bottom = y == J->th - 1 || J->parent[t + tw] == -1;
    const int gi = (J->ti0 + x) * kTile, gj = (J->tj0 + y) * kTile;
    ci0 = gi + (int)tile_field(s, 0);
    ci1 = gi + (int)tile_field(s, 1);
    cj0 = gj + (int)tile_field(s, 2);
    cj1 = gj + (int)tile_field(s, 3);
    first = cj0 * J->res + gi + (int)tile_field(s, 4);
  }
  // Every lane of a wavefront belongs to the same job (blocks never span
  // two jobs).
  wave_extreme(bodies, nslots, J->slot, left, id, 0, ci0, false);
  wave_extreme(bodies, nslots, J->slot, right, id, 1, ci1, true);
  wave_extreme(bodies, nslots, J->slot, top, id, 2, cj0, false);
  wave_extreme(bodies, nslots, J->slot, bottom, id, 3, cj1, true);
  wave_extreme(bodies, nslots, J->slot, top, id, 4, first, false);
}

// awake_work_boxes' tile test (src/fluid_core.rs:1537-1600) for a 2 cm
// sheet: one wavefront per 8 x 8 tile, one lane per cell.  The speed test
// is |(u, v)| > 0.05 with the magnitude formed as the host's hypotf forms it
// (single-precision result of the double-precision root of the double sum
// of squares); the tile also keeps its largest u^2 + v^2 (NaN never wins,
// as in the host's scan) and whether it holds a wet cell.
__global__ __launch_bounds__(64) void k_awake_tiles(const float* __restrict__ h, const float* __restrict__ u,
                                                    const float* __restrict__ v, const float* __restrict__ bed,
                                                    int res, float cell, float eps, unsigned char* hot,
                                                    unsigned* s2_bits, unsigned char* wet_any) {
  const int tiles = (res + 7) / 8;
  const int t = (int)blockIdx.x;
  const int ti = t % tiles, tj = t / tiles;
  const int lane = (int)threadIdx.x;
  const int i = ti * 8 + (lane & 7), j = tj * 8 + (lane >> 3);
  bool awake = false, wet = false;
  float s2 = 0.0f;
  if (i < res && j < res) {
    const size_t idx = (size_t)j * res + i;
    const float uu = u[idx], vv = v[idx];
    const float q = uu * uu + vv * vv;
    s2 = q > 0.0f ? q : 0.0f;
    const float hv = h[idx];
    wet = hv > eps;
    if (wet) {
      const float speed = (float)sqrt((double)uu * (double)uu + (double)vv * (double)vv);
      const float eta = bed[idx] + hv;
      const size_t ip = i + 1 < res ? idx + 1 : idx;
      const size_t jp = j + 1 < res ? idx + res : idx;
      const float dex = fabsf(bed[ip] + h[ip] - eta);
      const float dez = fabsf(bed[jp] + h[jp] - eta);
      awake = speed > 0.05f || fmaxf(dex, dez) > 0.02f * cell;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const float w = __shfl_xor(s2, o);
    s2 = w > s2 ? w : s2;
  }
  const bool any_awake = __ballot(awake) != 0ull, any_wet = __ballot(wet) != 0ull;
  if (lane == 0) {
    hot[t] = any_awake ? 1 : 0;
    wet_any[t] = any_wet ? 1 : 0;
    s2_bits[t] = __float_as_uint(s2);
  }
}

__global__ void k_zero_u32(unsigned* p, int n) {
  int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k < n) p[k] = 0u;
}

Box halo3_of(const SheetStats& st, int res) {
  return Box{std::max(st.i0 - 3, 0), std::min(st.i1 + 3, res - 1), std::max(st.j0 - 3, 0),
             std::min(st.j1 + 3, res - 1)};
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

int plan_boxes_many(PlanLabeller& L, const PlanSheet* s, int n, float dt, float eps, StepPlan* plans, char* host,
                    hipStream_t stream) {
  std::vector<int> lab;
  for (int k = 0; k < n; k++) {
    plans[k] = StepPlan{};
    host[k] = 0;
    const SheetStats& st = *s[k].stats;
    if (!(dt > 0.0f) || st.wet == 0) continue;
    if (s[k].cell <= kFineCell + 1.0e-6f) {  // 2 cm sheets plan on awake tiles
      host[k] = 1;
      continue;
    }
    const long long area = (long long)(st.i1 - st.i0 + 1) * (long long)(st.j1 - st.j0 + 1);
    if ((long long)st.wet == area) {
      plans[k].boxes.push_back(halo3_of(st, s[k].res));
      continue;
    }
    lab.push_back(k);
  }
  if (lab.empty()) return 0;
  const int nj = (int)lab.size();
  // Tile grids over each wet box, packed in one buffer.
  std::vector<PlanJob> jobs((size_t)nj);
  std::vector<size_t> off((size_t)nj);
  size_t tiles_bytes = 0;
  for (int q = 0; q < nj; q++) {
    const PlanSheet& sh = s[lab[q]];
    const SheetStats& st = *sh.stats;
    PlanJob& J = jobs[q];
    J.h = sh.h;
    J.res = sh.res;
    J.ti0 = st.i0 / kTile;
    J.tj0 = st.j0 / kTile;
    J.tw = st.i1 / kTile - J.ti0 + 1;
    J.th = st.j1 / kTile - J.tj0 + 1;
    J.bx = (J.tw + 63) / 64;
    J.cb = Box{st.i0, st.i1, st.j0, st.j1};
    J.eps = eps;
    J.slot = q;
    off[q] = tiles_bytes;
    tiles_bytes += align256((size_t)J.tw * J.th * sizeof(int)) * 2;
  }
  if (L.cap_tiles_bytes < tiles_bytes) {
    if (L.tiles) (void)hipFree(L.tiles);
    L.tiles = nullptr;
    L.cap_tiles_bytes = 0;
    if (hipMalloc(&L.tiles, tiles_bytes) != hipSuccess) return SILT_ENOMEM;
    L.cap_tiles_bytes = tiles_bytes;
  }
  for (int q = 0; q < nj; q++) {
    char* base = static_cast<char*>(L.tiles) + off[q];
    jobs[q].parent = reinterpret_cast<int*>(base);
    jobs[q].info = reinterpret_cast<unsigned*>(base + align256((size_t)jobs[q].tw * jobs[q].th * sizeof(int)));
  }
  if (L.cap_slots < nj) {
    const size_t body_ints = body_offset(nj, nj, kPlanBodyPrefix);  // counts, heads and tails of nj slots
    if (L.bodies) (void)hipFree(L.bodies);
    if (L.h_bodies) (void)hipHostFree(L.h_bodies);
    if (L.jobs) (void)hipFree(L.jobs);
    if (L.h_jobs) (void)hipHostFree(L.h_jobs);
    L.bodies = L.h_bodies = nullptr;
    L.jobs = L.h_jobs = nullptr;
    L.cap_slots = 0;
    if (hipMalloc(&L.bodies, body_ints * sizeof(int)) != hipSuccess ||
        hipHostMalloc(&L.h_bodies, body_ints * sizeof(int), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&L.jobs, (size_t)nj * sizeof(PlanJob)) != hipSuccess ||
        hipHostMalloc(&L.h_jobs, (size_t)nj * sizeof(PlanJob), hipHostMallocDefault) != hipSuccess)
      return SILT_ENOMEM;
    L.cap_slots = nj;
  }
  // The body layout depends on the slot count: lay it out for cap_slots.
  const int nslots = L.cap_slots;
  memcpy(L.h_jobs, jobs.data(), (size_t)nj * sizeof(PlanJob));
  if (hipMemcpyAsync(L.jobs, L.h_jobs, (size_t)nj * sizeof(PlanJob), hipMemcpyHostToDevice, stream) != hipSuccess)
    return SILT_EDEVICE;
  const PlanJob* dj = static_cast<const PlanJob*>(L.jobs);
  bool two_phase = false;
  for (const PlanJob& J : jobs) two_phase = two_phase || J.th > kJoinBlock;
  for (int c0 = 0; c0 < nj; c0 += kLaunchJobs) {
    const int c1 = std::min(nj, c0 + kLaunchJobs);
    JobStarts tiles{}, rows{};
    tiles.first = rows.first = c0;
    tiles.n = rows.n = c1 - c0;
    for (int q = c0; q < c1; q++) {
      const PlanJob& J = jobs[q];
      tiles.start[q - c0 + 1] = tiles.start[q - c0] + J.bx * ((J.th + 3) / 4);
      rows.start[q - c0 + 1] = rows.start[q - c0] + (J.th + 3) / 4;
    }
    const dim3 blk(64, 4);
    const unsigned gt = (unsigned)tiles.start[c1 - c0], gr = (unsigned)rows.start[c1 - c0];
    hipLaunchKernelGGL(k_tile_summary, dim3(gt), blk, 0, stream, tiles, dj, L.bodies);
    hipLaunchKernelGGL(k_tile_runs, dim3(gr), dim3(256), 0, stream, rows, dj);
    hipLaunchKernelGGL(k_tile_join, dim3(gt), blk, 0, stream, tiles, dj, 0);
    if (two_phase) {
      hipLaunchKernelGGL(k_tile_heads, dim3(gt), blk, 0, stream, tiles, dj);
      hipLaunchKernelGGL(k_tile_join, dim3(gt), blk, 0, stream, tiles, dj, 1);
    }
    hipLaunchKernelGGL(k_tile_heads, dim3(gt), blk, 0, stream, tiles, dj);
    hipLaunchKernelGGL(k_tile_cells, dim3(gt), blk, 0, stream, tiles, dj);
    hipLaunchKernelGGL(k_tile_roots, dim3(gt), blk, 0, stream, tiles, dj, L.bodies, nslots);
    hipLaunchKernelGGL(k_tile_boxes, dim3(gt), blk, 0, stream, tiles, dj, L.bodies, nslots);
  }
  // Counts and the first kPlanBodyPrefix records of every sheet: one copy.
  const size_t head_ints = body_offset(nslots, nj, 0);
  if (hipMemcpyAsync(L.h_bodies, L.bodies, head_ints * sizeof(int), hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return SILT_EDEVICE;
  const unsigned* count = reinterpret_cast<const unsigned*>(L.h_bodies);
  bool tail = false;
  for (int q = 0; q < nj; q++) {
    if (count[q] > (unsigned)kPlanBodyCap) continue;
    if (count[q] > (unsigned)kPlanBodyPrefix) {
      const size_t o = body_offset(nslots, q, kPlanBodyPrefix);
      if (hipMemcpyAsync(L.h_bodies + o, L.bodies + o,
                         (size_t)(count[q] - kPlanBodyPrefix) * kBodyStride * sizeof(int), hipMemcpyDeviceToHost,
                         stream) != hipSuccess)
        return SILT_EDEVICE;
      tail = true;
    }
  }
  if (tail && hipStreamSynchronize(stream) != hipSuccess) return SILT_EDEVICE;
  for (int q = 0; q < nj; q++) {
    const int k = lab[q];
    if (count[q] > (unsigned)kPlanBodyCap) {  // too many clusters to read back: plan on the host
      host[k] = 1;
      continue;
    }
    // Cluster numbers come from atomics; list them in the row-major order of
    // their first cell, as the reference lists its bodies.
    std::vector<std::pair<int, Box>> found((size_t)count[q]);
    for (unsigned id = 0; id < count[q]; id++) {
      const int* b = L.h_bodies + body_offset(nslots, q, (int)id);
      found[id] = {b[4], Box{b[0], b[1], b[2], b[3]}};
    }
    std::sort(found.begin(), found.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    std::vector<Box> bodies;
    bodies.reserve(found.size());
    for (auto& f : found) bodies.push_back(f.second);
    plans[k].boxes = boxes_from_bodies(std::move(bodies), s[k].res, halo3_of(*s[k].stats, s[k].res));
  }
  return 0;
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

int plan_fine_device(const float* h, const float* u, const float* v, const float* bed, int res, float cell, float dt,
                     float eps, float sub_cap, int pace, FineScratch& fs, StepPlan* plan, hipStream_t stream) {
  *plan = StepPlan{};
  if (!(dt > 0.0f) || res <= 0) return 1;
  if (pace == 2 || pace == 3) return 0;  // the whole-wet plans scan the planes on the host
  const int tiles = (res + 7) / 8, nt = tiles * tiles;
  if (fs.cap_tiles < nt) {
    if (fs.d) (void)hipFree(fs.d);
    if (fs.h) (void)hipHostFree(fs.h);
    fs.d = fs.h = nullptr;
    fs.cap_tiles = 0;
    if (hipMalloc(&fs.d, (size_t)nt * 6) != hipSuccess || hipHostMalloc(&fs.h, (size_t)nt * 6, 0) != hipSuccess)
      return SILT_ENOMEM;
    fs.cap_tiles = nt;
  }
  // Layout: s2 bits [nt] u32, hot [nt] u8, wet [nt] u8.
  unsigned* d_s2 = reinterpret_cast<unsigned*>(fs.d);
  unsigned char* d_hot = reinterpret_cast<unsigned char*>(fs.d) + (size_t)nt * 4;
  unsigned char* d_wet = d_hot + nt;
  hipLaunchKernelGGL(k_awake_tiles, dim3((unsigned)nt), dim3(64), 0, stream, h, u, v, bed, res, cell, eps, d_hot, d_s2,
                     d_wet);
  if (hipMemcpyAsync(fs.h, fs.d, (size_t)nt * 6, hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return SILT_EDEVICE;
  const unsigned* s2 = reinterpret_cast<const unsigned*>(fs.h);
  const unsigned char* hot = reinterpret_cast<const unsigned char*>(fs.h) + (size_t)nt * 4;
  const unsigned char* wet = hot + nt;
  // plan_step's wet-sum test: a sum of depths above eps exceeds eps exactly
  // when one cell is above eps.
  bool any = false;
  for (int k = 0; k < nt && !any; k++) any = wet[k] != 0;
  if (!any) return 1;
  plan->fine_open = pace == 0;
  plan->boxes = awake_boxes_from_hot(std::vector<uint8_t>(hot, hot + nt), res);
  if (plan->boxes.empty()) return 1;
  if (plan->fine_open) {
    plan->nsub = 1;
    return 1;
  }
  // The boxes are whole tiles (clipped to the grid): their speed maximum is
  // the maximum of their tiles'.
  float max_s2 = 0.0f;
  for (const Box& b : plan->boxes)
    for (int tj = b.j0 / 8; tj <= b.j1 / 8; tj++)
      for (int ti = b.i0 / 8; ti <= b.i1 / 8; ti++) {
        float q;
        memcpy(&q, &s2[tj * tiles + ti], 4);
        if (q > max_s2) max_s2 = q;
      }
  const float max_s = fmaxf(0.15f, sqrtf(max_s2));
  const float cap = std::isfinite(sub_cap) ? clampf(sub_cap, 1.0f, 8.0f) : 8.0f;
  const float q = ceilf((dt * max_s) / (0.42f * fmaxf(cell, 1.0e-3f)));
  plan->nsub = (int)clampf(q, 1.0f, cap);
  return 1;
}

void fine_scratch_free(FineScratch& fs) {
  if (fs.d) (void)hipFree(fs.d);
  if (fs.h) (void)hipHostFree(fs.h);
  fs = FineScratch{};
}

void plan_scratch_free(PlanScratch& ps) {
  if (ps.maxs2) (void)hipFree(ps.maxs2);
  if (ps.h_maxs2) (void)hipHostFree(ps.h_maxs2);
  ps = PlanScratch{};
}

void plan_labeller_free(PlanLabeller& L) {
  if (L.tiles) (void)hipFree(L.tiles);
  if (L.bodies) (void)hipFree(L.bodies);
  if (L.h_bodies) (void)hipHostFree(L.h_bodies);
  if (L.jobs) (void)hipFree(L.jobs);
  if (L.h_jobs) (void)hipHostFree(L.h_jobs);
  L = PlanLabeller{};
}

}  // namespace silt
