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

// Substep count of a gravity sheet whose boxes' largest squared speed is s2.
int nsub_from_s2(float s2, float dt, float cell);

// Device planning of gravity sheets with cell > 2 cm whose planes are on the
// device, from statistics of those planes: one box for a solid wet rectangle,
// else the wet set's clusters are labelled on the GPU (sheet_work_boxes).
// A sheet returns at most kPlanBodyCap clusters; the first kPlanBodyPrefix
// of every sheet come back with the counts in one copy.
constexpr int kPlanBodyCap = 4096;
constexpr int kPlanBodyPrefix = 64;

// Scratch of a labelling batch (grown on demand, freed with
// plan_labeller_free): tile labels and summaries, cluster records and their
// pinned copy, the sheet table.
struct PlanLabeller {
  void* tiles = nullptr;
  size_t cap_tiles_bytes = 0;
  int* bodies = nullptr;
  int* h_bodies = nullptr;  // pinned
  void* jobs = nullptr;
  void* h_jobs = nullptr;  // pinned
  int cap_slots = 0;
};
// One sheet to plan: depth plane on the device and its statistics.
struct PlanSheet {
  const float* h;
  int res;
  float cell;
  const SheetStats* stats;
};
// Work boxes of n sheets with one batched labelling and one wait.  plans[k]
// gets the boxes (nsub left 0); host[k] = 1 when sheet k must plan from its
// planes on the host (2 cm cell, or more than kPlanBodyCap clusters).
// Returns 0 or a SILT_E* code.
int plan_boxes_many(PlanLabeller& L, const PlanSheet* s, int n, float dt, float eps, StepPlan* plans, char* host,
                    hipStream_t stream);
void plan_labeller_free(PlanLabeller& L);

// Substep count of a planned sheet, split so many sheets wait once:
//   plan_nsub_begin  enqueue the boxes' speed maximum when the global one
//                    allows more than one substep;
//   (wait)           plan_nsub_end sets nsub.
struct PlanScratch {
  unsigned* maxs2 = nullptr;
  unsigned* h_maxs2 = nullptr;  // pinned
  bool maxs2_pending = false;   // plan_nsub_begin enqueued a box maximum
};
int plan_nsub_begin(const float* u, const float* v, int res, float cell, float dt, const SheetStats& st,
                    PlanScratch& ps, StepPlan* plan, hipStream_t stream);
void plan_nsub_end(PlanScratch& ps, float cell, float dt, StepPlan* plan);
void plan_scratch_free(PlanScratch& ps);

// Plan of a 2 cm sheet on the GPU for paces 0 and 1 (awake_work_boxes'
// tiles tested on the device, their flags and speed maxima read back, the
// boxes and substep count made on the host as plan_step makes them).
// Returns 1 when planned, 0 when the host must plan (paces 2 and 3), or a
// SILT_E* code.
struct FineScratch {
  void* d = nullptr;
  void* h = nullptr;  // pinned
  int cap_tiles = 0;
};
int plan_fine_device(const float* h, const float* u, const float* v, const float* bed, int res, float cell, float dt,
                     float eps, float sub_cap, int pace, FineScratch& fs, StepPlan* plan, hipStream_t stream);
void fine_scratch_free(FineScratch& fs);

// XCD balance of a launch: the 8 XCDs take workgroups round robin and each
// runs a contiguous run of tiles; weights[x] scales the rows of the tiles XCD
// x runs (every strip's rows are split in proportion to the weights of the
// XCDs that run its chunks), so a slower XCD gets shorter tiles.  clocks, when
// set, receives 4 values per workgroup: real-time clock at start and end
// (100 MHz), the rows its pipeline ran (XCC_ID register in the high word), and
// its tile (HW_ID register in the high word).
struct FusedBalance {
  const float* xcd_w = nullptr;  // [8], or null for uniform chunks
  unsigned long long* clocks = nullptr;
  // When set, the launch carries these events in its own dispatch
  // (hipExtLaunchKernelGGL): they time the kernel alone.
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
};
constexpr int kXcds = 8;
// The XCD that runs linear tile `tile` of a launch of `tiles` workgroups.
int fused_tile_xcd(int tile, int tiles);

// One substep of one or two layers (same res, drive kind and projection) in
// a single launch.  exact selects the reference-order arithmetic.
int fused_substep_layers(const FusedLayer* layers, int nlayers, const float* bed, int res, bool exact,
                         hipStream_t st, const FusedBalance* bal = nullptr);
// One substep of n independent windows (each with its own side, bed and
// planes; one drive kind and projection) in a single launch.  rows > 0 pins
// the rows per workgroup.
int fused_substep_windows(const FusedLayer* wins, int n, bool exact, int rows, hipStream_t st,
                          const FusedBalance* bal = nullptr);

// A prepared launch over n windows: the packed window descriptors (device
// layout, fused_desc_size() bytes each) and, for more than two windows, the
// window of every tile.  A caller that keeps its own device copies (a batch
// replayed from a graph) uploads descs and tile_window and enqueues.
struct FusedLaunch {
  bool exact = false, full = false, project = false, wind = false;
  int chunk = 0, nwin = 0, tiles = 0;
  std::vector<unsigned char> descs;
  std::vector<int> tile_window;
  std::vector<int> rows;  // [2 tiles] output rows of every tile, or empty (uniform chunks)
};
size_t fused_desc_size();
// Whether the fused kernel covers a box list: 1 to 8 boxes that neither
// overlap nor touch (so two cells in boxes are in one box exactly when a
// face between them exists).
bool fused_boxes_fit(const std::vector<Box>& boxes);
int fused_build(const FusedLayer* wins, int n, bool exact, int rows, FusedLaunch* out,
                const float* xcd_w = nullptr);
int fused_enqueue(const FusedLaunch& l, const void* dev_table, const int* dev_tile_window, hipStream_t st,
                  const int* dev_rows = nullptr, unsigned long long* clocks = nullptr,
                  hipEvent_t ev_start = nullptr, hipEvent_t ev_stop = nullptr);

// One planned step of one layer on the fused path.  Returns the substep count,
// SILT_EUNSUPPORTED for plans the fused kernel does not cover (pace-3 slices,
// touching or overlapping boxes, dry-bank transport), or another SILT_E* code.
int fused_step(const SheetPlanes& s, const StepPlan& plan, int res, float cell, float dt, const silt_drive& d,
               const silt_sheet_opts& o, bool gate_uniform, bool nu_uniform, float nu0, hipStream_t st,
               SheetTotals* book = nullptr);

// Runs a planned step of one layer on the path opts.path names (fused paths
// fall back to the stage kernels for plans they do not cover).
int run_plan(const SheetPlanes& s, const StepPlan& plan, int res, float cell, float dt, const silt_drive& d,
             const silt_sheet_opts& o, bool gate_uniform, bool nu_uniform, float nu0, hipStream_t st,
             SheetTotals* book = nullptr);

// A step of one layer on device planes, planned on the host from a read
// back of depth, velocity and bed (blocking); kernels enqueued on st.
int step_device_planes(const SheetPlanes& s, int res, float cell, float dt, const silt_drive& d,
                       const silt_sheet_opts& o, bool gate_uniform, bool nu_uniform, float nu0, hipStream_t st,
                       SheetTotals* book = nullptr);

// Output rows per workgroup: the smallest chunk that keeps every tile of the
// launch resident, unless fused_set_chunk_rows pinned one (0 unpins).
int fused_chunk_rows(int res, int strips, int layers, int resident);
void fused_set_chunk_rows(int rows);
// Pin the rows per workgroup for launches of `layers` sheets of side res
// (0: one- and two-sheet launches alike; rows 0 unpins); the pinned value,
// or 0 (layers 0: the two-sheet pin if any, else the one-sheet pin).
void fused_set_chunk_rows_for(int res, int rows, int layers = 0);
int fused_pinned_chunk_rows(int res, int layers = 0);

}  // namespace silt
