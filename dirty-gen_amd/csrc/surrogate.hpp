// Data and inference kernels of the tire sub-grid network (surrogate.hip).
//
// The network stands in for the nested 2 cm patch of a FluidCore step.  Its
// region R is the 16 x 16 coarse cells the patch covers grown by 2 (so up to
// 20 x 20, clipped to the window).  For every cell of R it reads a 16 x 16
// stencil of the coarse state before the step (offsets -8..+7, 6 channels)
// and predicts the volume that leaves the cell through each of its 4 edges,
// per layer, on top of a coarse-only step.  A face's flux is the mean of its
// two cells' predictions, so whatever one cell loses its neighbour gains and
// the volume balance closes by construction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "core.hpp"

namespace silt {

struct SheetTotals;
struct SheetStats;

constexpr int kSurReach = 8;                     // stencil offsets -8 .. +7
constexpr int kSurMax = 20;                      // largest region side
constexpr int kSurCtx = kSurMax + 2 * kSurReach;  // 36: context side (offsets -8..+8, for the mirrored stencils)
constexpr int kSurChannels = 7;                  // 6 inputs + inside-window flag
constexpr size_t kSurCtxFloats = (size_t)kSurChannels * kSurCtx * kSurCtx;
constexpr size_t kSurTgtFloats = (size_t)kSurMax * kSurMax * 8;
// Training targets are in millimetres of depth per step.
constexpr float kSurTargetScale = 1000.0f;

struct SurRegion {
  int i0, j0, nx, nz;  // grid cells [i0, i0 + nx) x [j0, j0 + nz)
};

// The coarse sheet the network reads and corrects: a res x res grid of
// cell metres with its (0, 0) corner at (ox, oz).
struct SurGrid {
  const float *water, *slurry, *uw, *vw, *bed;
  int res;
  float cell, ox, oz;
};

// The network's region for a group of pushes (one vehicle): the 2 cm
// patch's placement (refine, src/fluid_core.rs:8316-8356: 160 fine cells of
// 2 cm centred on the pushes' mean, snapped to the coarse grid) grown by 2
// cells, clipped to the grid.  nx = 0 when nothing of it is inside.
SurRegion sur_region_at(const CorePush* p, int n, float ox, float oz, float cell, int res);

// Context of one step: channels bed, water, slurry, water u and v relative to
// the tires' mean velocity (zero on dry cells), tire pressure, inside flag;
// each kSurCtx^2, cell (a, b) = grid cell (r.i0 - 8 + a, r.j0 - 8 + b).
void sur_context(const SurGrid& g, const SurRegion& r, const PushList& pl, float* ctx, hipStream_t st);
// Network inputs of every cell of R (row-major over R) from one context.
void sur_gather_region(const float* ctx, const SurRegion& r, uint16_t* xb, hipStream_t st);
// Random training samples: sample b draws a recorded step, a cell of its
// region and one of the square's 8 rotations / reflections from (seed, b);
// inputs (stencil, and the velocity vector, turned) into xb, targets (edges
// permuted to match) into tg.  The sheet's physics has the square's symmetry
// up to its x-then-z transport splitting, so the turned samples are the
// network's cheapest extra data.
void sur_gather_random(const float* ctx, const float* tgt, const SurRegion* regions, int nsteps, uint32_t seed,
                       int batch, uint16_t* xb, float* tg, hipStream_t st);
// Targets of one step: the edge fluxes that carry the coarse-only depths to
// the nested ones over R (least-norm, zero flux through R's border; the
// region's mean difference, spray and dry-cut differences, is left out),
// scaled by kSurTargetScale, [cell][E, W, N, S of water, E, W, N, S of slurry].
void sur_targets(const float* hn_w, const float* hc_w, const float* hn_s, const float* hc_s, int res,
                 const SurRegion& r, float* tgt, hipStream_t st);
// Apply predicted fluxes y ([cell][8], scaled) to the two depth planes:
// faces inside R carry the mean of their two cells' predictions, limited so
// no donor gives more than it holds.  fac: scratch of 2 * nx * nz floats.
// totals (water, slurry), when given, book in FP64 what the FP32 update of
// each cell adds beyond its exchanged fluxes (their residue line); stats
// (water, slurry), when given, get the change of their FP64 volume (the sum
// of max(h, 0)), so the volume stays that of the planes (the other fields
// do not follow).
void sur_apply(float* water, float* slurry, int res, const float* y, const SurRegion& r, float* fac,
               hipStream_t st, SheetTotals* totals = nullptr, SheetStats* stats = nullptr);
// Many vehicles in one launch per stage: vehicle v's region, its first row
// of network inputs / outputs, its pushes and their mean velocity, and the
// box its pushes shove.  Regions (and push boxes) must not overlap.  The
// contexts of vehicle v sit at ctx + v * kSurCtxFloats, its factor scratch
// at fac + v * 2 * kSurMax^2.
struct TireJob {
  SurRegion r;
  int row;
  float vtx, vtz;
  PushBox push;
};
// Mean velocity of pushes p[0..n) along their headings.
void sur_mean_velocity(const CorePush* p, int n, float* vx, float* vz);
void sur_context_many(const SurGrid& g, const TireJob* jobs, const CorePush* pushes, int n, float* ctx,
                      hipStream_t st);
void sur_gather_many(const float* ctx, const TireJob* jobs, int n, uint16_t* xb, hipStream_t st);
void sur_apply_many(float* water, float* slurry, int res, const float* y, const TireJob* jobs, int n, float* fac,
                    hipStream_t st, SheetTotals* totals = nullptr, SheetStats* stats = nullptr);
// Sum over the region of (a - b)^2 added into *out (FP64).
void sur_region_sq_diff(const float* a, const float* b, int res, const SurRegion& r, double* out, hipStream_t st);
// Largest |a - b| over n cells, as float bits, max-ed into *out.
void sur_max_abs_diff(const float* a, const float* b, int n, unsigned* out, hipStream_t st);

}  // namespace silt
