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
  SheetPlanes in;  // in.bed is the window's bed
  int res = 0;     // side of the window (fused_substep_windows)
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

// Plan of a gravity sheet from its statistics; false when the wet set is not
// one solid rectangle (or the cell is 2 cm) and the planes must be read back.
// need_box_max: nsub > 1 from the global maximum speed, to be refined with the
// box's own maximum (box_max_s2).
bool plan_from_stats(const SheetStats& st, int res, float cell, float dt, StepPlan* plan, bool* need_box_max);
// Substep count of a gravity sheet whose boxes' largest squared speed is s2.
int nsub_from_s2(float s2, float dt, float cell);

// One substep of one or two layers (same res, drive kind and projection) in
// a single launch.  exact selects the reference-order arithmetic.
int fused_substep_layers(const FusedLayer* layers, int nlayers, const float* bed, int res, bool exact,
                         hipStream_t st);
// One substep of n independent windows (each with its own side, bed and
// planes; one drive kind and projection) in a single launch.  rows > 0 pins
// the rows per workgroup.
int fused_substep_windows(const FusedLayer* wins, int n, bool exact, int rows, hipStream_t st);

// A prepared launch over n windows: the packed window descriptors (device
// layout, fused_desc_size() bytes each) and, for more than two windows, the
// window of every tile.  A caller that keeps its own device copies (a batch
// replayed from a graph) uploads descs and tile_window and enqueues.
struct FusedLaunch {
  bool exact = false, full = false, project = false, wind = false;
  int chunk = 0, nwin = 0, tiles = 0;
  std::vector<unsigned char> descs;
  std::vector<int> tile_window;
};
size_t fused_desc_size();
int fused_build(const FusedLayer* wins, int n, bool exact, int rows, FusedLaunch* out);
int fused_enqueue(const FusedLaunch& l, const void* dev_table, const int* dev_tile_window, hipStream_t st);

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
