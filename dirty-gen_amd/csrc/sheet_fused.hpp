// Fused streaming sheet step (sheet_fused.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "plan.hpp"
#include "sheet_device.hpp"
#include "silt_sheet.h"

namespace silt {

// Statistics of one layer's planes that the next step's plan reads: FP64
// volume sum (metres of depth, times cell area on the host), wet count and
// wet bounding box (h > eps), largest squared speed.  Zero it with
// stats_reset before a kernel accumulates into it.
struct SheetStats {
  double vol;
  unsigned long long wet;
  int i0, i1, j0, j1;
  unsigned max_s2_bits;
  int pad;
};

// Running FP64 ledger lines of one layer, metres of depth summed over cells:
// depth the dry cut deleted, and the transport residue (FP32 rounding and the
// max(0) clamp beyond the exchanged fluxes).
struct SheetTotals {
  double cut;
  double residue;
};

// One layer of one fused substep: read planes in `in`, write ho/uo/vo.
struct FusedLayer {
  SheetPlanes in;
  float* ho;
  float* uo;
  float* vo;
  StepArgs a;
  std::vector<Box> boxes;
  SheetStats* stats = nullptr;
  SheetTotals* totals = nullptr;
};

// Zeroes n SheetStats (bounds set empty) with a kernel, never a memset, so
// the call can sit inside a captured graph.
void stats_reset(SheetStats* s, int n, hipStream_t st);
// Stats of planes (h, u, v) from scratch: one read pass.
void stats_compute(const float* h, const float* u, const float* v, int res, float eps, SheetStats* s,
                   hipStream_t st);
// Largest u^2 + v^2 over the cells of the given boxes, as float bits in *out
// (out zeroed first by the caller's stats_reset or a kernel).
void box_max_s2(const float* u, const float* v, int res, const Box& b, unsigned* out, hipStream_t st);

StepArgs make_args(int res, float cell, float dt, const silt_drive& d, float eps, int banks, bool gate_uniform,
                   bool nu_uniform, float nu0);

// One substep of one or two layers (same res, drive kind and projection) in
// a single launch.  exact selects the reference-order arithmetic.
int fused_substep_layers(const FusedLayer* layers, int nlayers, const float* bed, int res, bool exact,
                         hipStream_t st);

// One planned step of one layer on the fused path.  Returns the substep count,
// SILT_EUNSUPPORTED for plans the fused kernel does not cover (pace-3 slices,
// touching or overlapping boxes, dry-bank transport), or another SILT_E* code.
int fused_step(const SheetPlanes& s, const StepPlan& plan, int res, float cell, float dt, const silt_drive& d,
               const silt_sheet_opts& o, bool gate_uniform, bool nu_uniform, float nu0, hipStream_t st);

// Runs a planned step of one layer on the path opts.path names (fused paths
// fall back to the stage kernels for plans they do not cover).
int run_plan(const SheetPlanes& s, const StepPlan& plan, int res, float cell, float dt, const silt_drive& d,
             const silt_sheet_opts& o, bool gate_uniform, bool nu_uniform, float nu0, hipStream_t st);

// A step of one layer on device planes, planned on the host from a read
// back of depth, velocity and bed (blocking); kernels enqueued on st.
int step_device_planes(const SheetPlanes& s, int res, float cell, float dt, const silt_drive& d,
                       const silt_sheet_opts& o, bool gate_uniform, bool nu_uniform, float nu0, hipStream_t st);

// Output rows per workgroup: the smallest chunk that keeps every tile of the
// launch resident, unless fused_set_chunk_rows pinned one (0 unpins).
int fused_chunk_rows(int res, int strips, int layers, int resident);
void fused_set_chunk_rows(int rows);
// Pin the rows per workgroup for sheets of side res (0 unpins); the pinned
// value for res, or 0.
void fused_set_chunk_rows_for(int res, int rows);
int fused_pinned_chunk_rows(int res);

}  // namespace silt
