// The two-layer FluidCore on the GPU (include/silt_core.h).  The host keeps
// the small decisions of a step (where the 2 cm patch goes, which coarse
// columns it holds, which tires shove which grid); every plane operation is
// a kernel on the handle's stream and the planes never leave the device,
// except the read back the sheet step's work-box plan takes.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <unordered_map>
#include <vector>

#include "core.hpp"
#include "tracers.hpp"
#include "mlp.hpp"
#include "plan.hpp"
#include "pond.hpp"
#include "rut.hpp"
#include "sheet_device.hpp"
#include "sheet_fused.hpp"
#include "silt_core.h"
#include "silt_mlp.h"
#include "surrogate.hpp"
#include "workspace.hpp"

#define HIPCHK(x)                                 \
  do {                                            \
    if ((x) != hipSuccess) return SILT_EDEVICE;   \
  } while (0)

using namespace silt;

namespace {

constexpr int WN = kWindowRes * kWindowRes;
constexpr int FN = kFineRes * kFineRes;
constexpr unsigned kNearCap = 1u << 14;
constexpr unsigned kSprayCap = 1u << 16;

}  // namespace

struct silt_core {
  float origin[2] = {0, 0};
  float center[2] = {0, 0};  // chase point (set_center)
  float moisture_ambient = 0.0f;  // read_substrate's most common moisture
  // Coarse cells parked off the window by recenter, by world cell key.
  struct Parked {
    float water, slurry, uw, vw, us, vs, moisture, foam, turbid, susp;
  };
  std::unordered_map<uint64_t, Parked> parked;
  float fine_origin[2] = {0, 0};
  int path = SILT_PATH_EXACT;
  int fine_pace = 0;
  float fine_sub_cap = 8.0f;
  int coarse_confine = SILT_CONFINE_REFERENCE;
  bool near_on = true;
  bool fine_on = false;
  long steps = 0;
  std::vector<CorePush> pushes;
  std::vector<uint8_t> hold;  // host mirror of the device hold mask
  CorePlanes c{};
  FinePlanes f{};
  uint8_t* d_hold = nullptr;
  float* nu_c = nullptr;
  float* nu_f = nullptr;
  NearTable near[2]{};
  int near_cur = 0;
  NearColumn* d_cols = nullptr;
  SprayEvent* d_ev = nullptr;
  unsigned* d_nev = nullptr;
  double* d_acc = nullptr;  // [0] spray volume this step, [1..2] frozen water / slurry
  double spray_m3 = 0.0;
  // FP64 ledger lines (CoreBook) and, behind them, the copy the judged
  // surrogate step restores.  base: each layer's total at the last caller
  // edit less the lines booked by then; rebase: an edit since.
  double* d_book = nullptr;
  double base[2] = {0.0, 0.0};
  bool rebase = true;
  // Hydro solids (set_hydro_solids), their mask descriptors, the wall masks
  // of the window and the patch, and the scour's bed drop, eroded volume
  // (FP64, m^3) and deepest drop (float bits).
  std::vector<silt_hydro_solid> solids;
  SolidDesc* d_solids = nullptr;
  uint8_t* d_mask = nullptr;
  uint8_t* d_fmask = nullptr;
  float* d_bed_drop = nullptr;
  double* d_eroded = nullptr;
  unsigned* d_hole = nullptr;
  // Foam and stirred-silt tracers of the window, their pass control and the
  // last step's pushes (note_surface_foam reads them).
  TracerPlanes tr{};
  TracerState* d_tstate = nullptr;
  // The wheel-lifted suspended solids and their scratch.
  SuspPlanes sus{};
  // Scratch and result of pond_ruts / seat_rut_water.
  PondScratch pond{};
  PondResult* d_pond = nullptr;
  std::vector<CorePush> last_pushes;
  // Ground under the sheet (SoilField bed / placed / initial height on the
  // window grid), set by silt_core_set_ground and cut by the rut stamp.
  bool ground = false;
  float *soil_bed = nullptr, *soil_placed = nullptr, *soil_init = nullptr, *soil_rut = nullptr;
  RutContact* d_contacts = nullptr;
  // Tire model (SILT_TIRE_*) and the surrogate's state: the network, its
  // context and limiter scratch, the judge's snapshot of the six planes and
  // the surrogate's depths, and the events that time the tire block.
  int tire_model = SILT_TIRE_NESTED;
  bool net_ready = false;
  Mlp net;
  float* sur_ctx = nullptr;
  float* sur_fac = nullptr;
  float* snap[6] = {};
  float* judged_h[2] = {};
  unsigned* d_err = nullptr;
  double* d_sq = nullptr;
  int judge_every = 30;
  float judge_tol = 1.0e-3f;
  int fallback_left = 0;
  long sur_steps = 0;
  float last_judge_err = -1.0f;
  float last_judge_rms = -1.0f;
  int judged = 0, fallbacks = 0, last_mode = SILT_TIRE_NESTED;
  float last_tire_ms = 0.0f;
  // Tire block timing: events 0-1 bracket the patch (nested) or the
  // prediction (surrogate), 2-3 the surrogate's flux apply.
  hipEvent_t tev[4] = {nullptr, nullptr, nullptr, nullptr};
  int timed = 0;  // 0 none, 1 nested, 2 surrogate
  // The slurry layer's stream and the events that fence it to the main one.
  hipStream_t st2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  // Device planning of the coarse pair: statistics of both layers, the
  // batched wet-set labelling, the substep maxima.
  SheetStats* d_stats = nullptr;
  SheetStats* h_stats = nullptr;  // pinned
  // The ledger's statistics of the four planes (coarse and fine layers).
  SheetStats* d_lstats = nullptr;
  // Scratch of the publish calls: 10 output planes and 3 soil inputs.
  float* d_pub = nullptr;
  PlanLabeller lab;
  PlanScratch ps[2];
  FineScratch fs[2];  // the 2 cm sheets' tile tests, per layer
  hipStream_t st = nullptr;
  std::vector<void*> allocs;
};

namespace {

// Layout of silt_core::d_book: what each sheet step's transports and dry cuts
// booked (depth summed over cells of the 20 cm window and of the 2 cm patch,
// water and slurry), then the handoffs' FP64 volumes in m^3.
struct CoreBook {
  SheetTotals coarse[2];
  SheetTotals fine[2];
  double handoff[2];
  double source[2];  // volume that entered the layers from the bed (scour)
};
constexpr size_t kBookDoubles = sizeof(CoreBook) / sizeof(double);

SheetTotals* coarse_book(silt_core* c, int l) { return reinterpret_cast<CoreBook*>(c->d_book)->coarse + l; }
SheetTotals* fine_book(silt_core* c, int l) { return reinterpret_cast<CoreBook*>(c->d_book)->fine + l; }
double* handoff_book(silt_core* c) { return reinterpret_cast<CoreBook*>(c->d_book)->handoff; }
double* source_book(silt_core* c) { return reinterpret_cast<CoreBook*>(c->d_book)->source; }

template <typename T>
bool dalloc(silt_core* c, T** p, size_t n) {
  if (hipMalloc((void**)p, n * sizeof(T)) != hipSuccess) return false;
  c->allocs.push_back((void*)*p);
  return hipMemset(*p, 0, n * sizeof(T)) == hipSuccess;
}

bool near_alloc(silt_core* c, NearTable* t) {
  t->cap = kNearCap;
  return dalloc(c, &t->keys, kNearCap) && dalloc(c, &t->data, (size_t)kNearCap * 400) && dalloc(c, &t->peak, kNearCap) &&
         dalloc(c, &t->idle, kNearCap) && dalloc(c, &t->fresh, kNearCap) && dalloc(c, &t->count, 1);
}

// coarse_axis (src/fluid_core.rs:9557-9568); -1 off the window.
int coarse_axis(float origin, float p) {
  float i = floorf((p - origin) / kCoarseCell);
  if (i < 0.0f || i >= (float)kWindowRes) return -1;
  return (int)i;
}

void coarse_xz(const silt_core* c, int i, int j, float* x, float* z) {
  *x = c->origin[0] + ((float)i + 0.5f) * kCoarseCell;
  *z = c->origin[1] + ((float)j + 0.5f) * kCoarseCell;
}

FineAxes fine_axes(const silt_core* c) {
  FineAxes ax{};
  for (int i = 0; i < kFineRes; i++)
    ax.col[i] = (short)coarse_axis(c->origin[0], c->fine_origin[0] + ((float)i + 0.5f) * kFineCell);
  for (int j = 0; j < kFineRes; j++)
    ax.row[j] = (short)coarse_axis(c->origin[1], c->fine_origin[1] + ((float)j + 0.5f) * kFineCell);
  return ax;
}

bool fine_held_at(const silt_core* c, float x, float z) {
  int i = coarse_axis(c->origin[0], x), j = coarse_axis(c->origin[1], z);
  if (i < 0 || j < 0) return false;
  return c->hold[(size_t)j * kWindowRes + i] == 1;
}

// cell_is_refined (src/fluid_core.rs:8337-8356)
bool cell_is_refined(const silt_core* c, float x, float z) {
  if (!c->fine_on || !fine_held_at(c, x, z)) return false;
  int i = coarse_axis(c->origin[0], x), j = coarse_axis(c->origin[1], z);
  float x0 = c->origin[0] + (float)i * kCoarseCell, z0 = c->origin[1] + (float)j * kCoarseCell;
  float x1 = x0 + kCoarseCell, z1 = z0 + kCoarseCell;
  float extent = (float)kFineRes * kFineCell;
  return x0 >= c->fine_origin[0] - 1.0e-4f && z0 >= c->fine_origin[1] - 1.0e-4f &&
         x1 <= c->fine_origin[0] + extent + 1.0e-4f && z1 <= c->fine_origin[1] + extent + 1.0e-4f;
}

float dist2_segment(float px, float pz, float ax, float az, float bx, float bz) {
  float abx = bx - ax, abz = bz - az;
  float ab2 = abx * abx + abz * abz;
  float t = ab2 < 1.0e-8f ? 0.0f : clampf(((px - ax) * abx + (pz - az) * abz) / ab2, 0.0f, 1.0f);
  float dx = px - (ax + abx * t), dz = pz - (az + abz * t);
  return dx * dx + dz * dz;
}

// near_tire_zone (src/fluid_core.rs:9624-9646)
bool near_tire_zone(const std::vector<CorePush>& p, float x, float z, float r2) {
  for (const CorePush& q : p) {
    float dx = x - q.x, dz = z - q.z;
    if (dx * dx + dz * dz <= r2) return true;
  }
  for (size_t a = 0; a < p.size(); a++)
    for (size_t b = a + 1; b < p.size(); b++) {
      if (fabsf(p[a].x - p[b].x) >= 0.45f) continue;
      if (dist2_segment(x, z, p[a].x, p[a].z, p[b].x, p[b].z) <= r2) return true;
    }
  return false;
}

// Runs of fine cells that map to one coarse index (near_columns, near.rs:152-183).
std::vector<std::pair<int, int>> axis_runs(const short* axis) {
  std::vector<std::pair<int, int>> out;
  int i = 0;
  while (i < kFineRes) {
    int cc = axis[i], n = 1;
    while (i + n < kFineRes && axis[i + n] == cc) n++;
    if (cc != -1 && n == 10) out.push_back({cc, i});
    i += n;
  }
  return out;
}

std::vector<NearColumn> near_columns(const silt_core* c) {
  FineAxes ax = fine_axes(c);
  auto cols = axis_runs(ax.col), rows = axis_runs(ax.row);
  std::vector<NearColumn> out;
  for (auto& r : rows)
    for (auto& q : cols) {
      int cidx = r.first * kWindowRes + q.first;
      if (c->hold[cidx] != 1) continue;
      float x, z;
      coarse_xz(c, q.first, r.first, &x, &z);
      NearColumn nc{};
      nc.cidx = cidx;
      nc.kx = (int)floorf(x / kCoarseCell);
      nc.kz = (int)floorf(z / kCoarseCell);
      nc.i0 = q.second;
      nc.j0 = r.second;
      int i = q.first, j = r.first;
      nc.inner = (i > 0 && j > 0 && i + 1 < kWindowRes && j + 1 < kWindowRes && c->hold[cidx - 1] == 1 &&
                  c->hold[cidx + 1] == 1 && c->hold[cidx - kWindowRes] == 1 && c->hold[cidx + kWindowRes] == 1)
                     ? 1
                     : 0;
      out.push_back(nc);
    }
  return out;
}

int upload_cols(silt_core* c, const std::vector<NearColumn>& cols) {
  if (cols.empty()) return 0;
  HIPCHK(hipMemcpyAsync(c->d_cols, cols.data(), cols.size() * sizeof(NearColumn), hipMemcpyHostToDevice, c->st));
  return 0;
}

int near_fade(silt_core* c, float dt) {
  float d = (dt > 0.0f && isfinite(dt)) ? dt : 0.0f;
  core_near_fade(c->near[c->near_cur], d, c->st);
  return 0;
}

// Coarse cells under the patch with their fine column / row ranges.
CoarseSpans coarse_spans(const FineAxes& ax) {
  CoarseSpans sp{};
  int ci0 = 1 << 30, ci1 = -1, cj0 = 1 << 30, cj1 = -1;
  for (int i = 0; i < kFineRes; i++)
    if (ax.col[i] >= 0) ci0 = std::min(ci0, (int)ax.col[i]), ci1 = std::max(ci1, (int)ax.col[i]);
  for (int j = 0; j < kFineRes; j++)
    if (ax.row[j] >= 0) cj0 = std::min(cj0, (int)ax.row[j]), cj1 = std::max(cj1, (int)ax.row[j]);
  if (ci1 < 0 || cj1 < 0) return sp;
  sp.ci0 = ci0;
  sp.cj0 = cj0;
  sp.ncol = std::min(ci1 - ci0 + 1, 24);
  sp.nrow = std::min(cj1 - cj0 + 1, 24);
  for (int a = 0; a < sp.ncol; a++) {
    sp.col_start[a] = -1;
    for (int i = 0; i < kFineRes; i++)
      if (ax.col[i] == ci0 + a) {
        if (sp.col_start[a] < 0) sp.col_start[a] = (short)i;
        sp.col_count[a]++;
      }
    if (sp.col_start[a] < 0) sp.col_start[a] = 0;
  }
  for (int b = 0; b < sp.nrow; b++) {
    sp.row_start[b] = -1;
    for (int j = 0; j < kFineRes; j++)
      if (ax.row[j] == cj0 + b) {
        if (sp.row_start[b] < 0) sp.row_start[b] = (short)j;
        sp.row_count[b]++;
      }
    if (sp.row_start[b] < 0) sp.row_start[b] = 0;
  }
  return sp;
}

int restrict_fine(silt_core* c) {
  if (!c->fine_on) return 0;
  if (c->near_on) {
    auto cols = near_columns(c);
    if (upload_cols(c, cols)) return SILT_EDEVICE;
    core_near_capture(c->f, c->near[c->near_cur], c->d_cols, (int)cols.size(), c->st);
  }
  FineAxes ax = fine_axes(c);
  CoarseSpans sp = coarse_spans(ax);
  core_restrict(c->c, c->f, c->d_hold, sp, ax, c->d_acc + 1, handoff_book(c), c->st);
  core_fill(c->f.water, 0.0f, FN, c->st);
  core_fill(c->f.slurry, 0.0f, FN, c->st);
  c->fine_on = false;
  return 0;
}

int refine(silt_core* c, const std::vector<CorePush>& p, float dt) {
  int rc = restrict_fine(c);
  if (rc) return rc;
  if (p.empty()) return near_fade(c, dt);
  float cx = 0.0f, cz = 0.0f;
  for (const CorePush& q : p) {
    cx += q.x;
    cz += q.z;
  }
  cx /= (float)p.size();
  cz /= (float)p.size();
  float extent = (float)kFineRes * kFineCell;
  c->fine_origin[0] = floorf((cx - extent * 0.5f) / kCoarseCell) * kCoarseCell;
  c->fine_origin[1] = floorf((cz - extent * 0.5f) / kCoarseCell) * kCoarseCell;
  float* zero[6] = {c->f.water, c->f.slurry, c->f.uw, c->f.vw, c->f.us, c->f.vs};
  for (float* z : zero) core_fill(z, 0.0f, FN, c->st);
  std::fill(c->hold.begin(), c->hold.end(), 0);
  FineAxes ax = fine_axes(c);
  CoarseSpans sp = coarse_spans(ax);
  const float r2 = 0.56f * 0.56f;
  for (int b = 0; b < sp.nrow; b++)
    for (int a = 0; a < sp.ncol; a++) {
      if (sp.col_count[a] * sp.row_count[b] != 100) continue;
      int i = sp.ci0 + a, j = sp.cj0 + b;
      float x, z;
      coarse_xz(c, i, j, &x, &z);
      if (near_tire_zone(p, x, z, r2)) c->hold[(size_t)j * kWindowRes + i] = 1;
    }
  HIPCHK(hipMemcpyAsync(c->d_hold, c->hold.data(), WN, hipMemcpyHostToDevice, c->st));
  core_refine_copy(c->c, c->f, c->d_hold, ax, sp.ci0, sp.cj0, sp.ncol, sp.nrow, c->st);
  if (c->near_on) {
    auto cols = near_columns(c);
    if (upload_cols(c, cols)) return SILT_EDEVICE;
    core_near_apply(c->f, c->near[c->near_cur], c->d_cols, (int)cols.size(), handoff_book(c), c->st);
  }
  near_fade(c, dt);
  c->fine_on = true;
  return 0;
}

PushList push_list(const std::vector<CorePush>& v) {
  PushList pl{};
  pl.n = (int)std::min(v.size(), (size_t)kMaxPush);
  for (int k = 0; k < pl.n; k++) pl.p[k] = v[k];
  return pl;
}

// step_pair (src/fluid_core.rs:7942-8150): the water and slurry layers are
// independent (the reference joins them on two threads), so the slurry runs
// on the handle's second stream with its own scratch, fenced by events.
// The coarse pair (step_pair on the 200^2 window, src/fluid_core.rs:7942-8150)
// planned on the GPU: soften both layers, then the statistics of both,
// their boxes from one batched labelling and their substep counts (the
// same plan as the host scan: gravity, pace 0, cap 8), then each layer's
// stages on its own stream.
int step_coarse_pair(silt_core* c, float dt, Workspace* ws, Workspace* ws2, const silt_sheet_opts& o, float nu_w) {
  const int res = kWindowRes;
  const float cell = kCoarseCell;
  core_soften(c->c.uw, c->c.vw, res, c->st);
  float* nu = c->nu_c;
  core_nu(nu, c->c.grain, c->c.moisture, res * res, c->st2);
  core_soften(c->c.us, c->c.vs, res, c->st2);
  stats_compute(c->c.water, c->c.uw, c->c.vw, res, o.eps, c->d_stats, c->st);
  stats_compute(c->c.slurry, c->c.us, c->c.vs, res, o.eps, c->d_stats + 1, c->st2);
  HIPCHK(hipMemcpyAsync(c->h_stats, c->d_stats, sizeof(SheetStats), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(c->h_stats + 1, c->d_stats + 1, sizeof(SheetStats), hipMemcpyDeviceToHost, c->st2));
  HIPCHK(hipStreamSynchronize(c->st2));
  HIPCHK(hipStreamSynchronize(c->st));
  PlanSheet sheets[2] = {{c->c.water, res, cell, c->h_stats}, {c->c.slurry, res, cell, c->h_stats + 1}};
  StepPlan plans[2];
  char host[2];
  int rc = plan_boxes_many(c->lab, sheets, 2, dt, o.eps, plans, host, c->st);
  if (rc < 0) return rc;
  const float* uv[2][2] = {{c->c.uw, c->c.vw}, {c->c.us, c->c.vs}};
  bool wait = false;
  for (int l = 0; l < 2; l++) {
    if (host[l]) continue;
    if ((rc = plan_nsub_begin(uv[l][0], uv[l][1], res, cell, dt, c->h_stats[l], c->ps[l], &plans[l], c->st)) < 0)
      return rc;
    wait = wait || c->ps[l].maxs2_pending;
  }
  if (wait) HIPCHK(hipStreamSynchronize(c->st));
  for (int l = 0; l < 2; l++)
    if (!host[l]) plan_nsub_end(c->ps[l], cell, dt, &plans[l]);
  // Water on the main stream, slurry on the second (both wait for the
  // planning above, which read both layers on the main stream).
  HIPCHK(hipEventRecord(c->fork, c->st));
  HIPCHK(hipStreamWaitEvent(c->st2, c->fork, 0));
  SheetPlanes s = ws->planes();
  s.h = c->c.water;
  s.u = c->c.uw;
  s.v = c->c.vw;
  s.bed = c->c.bed;
  silt_drive dw{SILT_DRIVE_GRAVITY, 0.0f, 0.0f, 0.08f};
  rc = host[0] ? step_device_planes(s, res, cell, dt, dw, o, true, true, nu_w, c->st, coarse_book(c, 0))
               : (plans[0].boxes.empty()
                      ? 0
                      : run_plan(s, plans[0], res, cell, dt, dw, o, true, true, nu_w, c->st, coarse_book(c, 0)));
  if (rc < 0) return rc;
  SheetPlanes s2 = ws2->planes();
  s2.h = c->c.slurry;
  s2.u = c->c.us;
  s2.v = c->c.vs;
  s2.bed = c->c.bed;
  s2.nu = nu;
  silt_drive ds{SILT_DRIVE_GRAVITY, 0.0f, 0.0f, 0.05f};
  rc = host[1] ? step_device_planes(s2, res, cell, dt, ds, o, true, false, 0.0f, c->st2, coarse_book(c, 1))
               : (plans[1].boxes.empty()
                      ? 0
                      : run_plan(s2, plans[1], res, cell, dt, ds, o, true, false, 0.0f, c->st2, coarse_book(c, 1)));
  HIPCHK(hipEventRecord(c->join, c->st2));
  HIPCHK(hipStreamWaitEvent(c->st, c->join, 0));
  return rc < 0 ? rc : 0;
}

// The wall mask of one grid from its planes now (solid_mask,
// src/fluid_core.rs:7288-7362): descriptors on the host (the cell ranges and
// the yaw's sine and cosine, as the reference computes them), the per-cell
// test on the GPU.
int solid_mask_device(silt_core* c, bool fine, uint8_t* mask, hipStream_t st) {
  const int res = fine ? kFineRes : kWindowRes;
  const float cell = fine ? kFineCell : kCoarseCell;
  const float* org = fine ? c->fine_origin : c->origin;
  std::vector<SolidDesc> d;
  for (int pass = 0; pass < 2; pass++)
    for (const silt_hydro_solid& s : c->solids) {
      if ((s.contact != 0) != (pass == 1)) continue;
      if (s.half_length <= 1.0e-4f || s.half_width <= 1.0e-4f) continue;
      const float reach = hypotf(fmaxf(s.half_length, 0.0f), fmaxf(s.half_width, 0.0f)) + cell;
      if (!isfinite(reach) || !isfinite(s.x) || !isfinite(s.z)) continue;
      const float lo_i = floorf((s.x - reach - org[0]) / cell - 0.5f);
      const float hi_i = ceilf((s.x + reach - org[0]) / cell - 0.5f);
      const float lo_j = floorf((s.z - reach - org[1]) / cell - 0.5f);
      const float hi_j = ceilf((s.z + reach - org[1]) / cell - 0.5f);
      const float top = (float)(res - 1);
      if (hi_i < 0.0f || hi_j < 0.0f || lo_i > top || lo_j > top) continue;
      SolidDesc q{};
      q.x = s.x;
      q.z = s.z;
      q.fwd_x = sinf(s.yaw);
      q.fwd_z = cosf(s.yaw);
      q.half_length = s.half_length;
      q.half_width = s.half_width;
      q.bottom_y = s.bottom_y;
      q.top_y = s.top_y;
      q.i0 = (int)fmaxf(lo_i, 0.0f);
      q.i1 = (int)fminf(hi_i, top);
      q.j0 = (int)fmaxf(lo_j, 0.0f);
      q.j1 = (int)fminf(hi_j, top);
      q.value = pass == 1 ? 1 : 2;
      d.push_back(q);
    }
  if (!d.empty()) HIPCHK(hipMemcpyAsync(c->d_solids, d.data(), d.size() * sizeof(SolidDesc), hipMemcpyHostToDevice, st));
  const float* water = fine ? c->f.water : c->c.water;
  const float* bed = fine ? c->f.bed : c->c.bed;
  core_solid_mask(water, bed, res, cell, org[0], org[1], c->d_solids, (int)d.size(), mask, st);
  // The descriptors are read by the launch above before the next call
  // rewrites them: the copy is ordered on the same stream.
  return hipGetLastError() == hipSuccess ? 0 : SILT_EDEVICE;
}

// step_pair with walls (src/fluid_core.rs:7954-8125): the grid's mask from
// the planes before either layer steps, then each layer softened (window
// only), displaced out of the walls and stepped as a walled sheet, water on
// the main stream and slurry on the second.  Both plan on the host.
int step_pair_walled(silt_core* c, float dt, bool fine) {
  const int res = fine ? kFineRes : kWindowRes;
  const float cell = fine ? kFineCell : kCoarseCell;
  Workspace* ws = workspace_for(res, 0);
  Workspace* ws2 = workspace_for(res, 1);
  if (!ws || !ws2) return SILT_ENOMEM;
  uint8_t* mask = fine ? c->d_fmask : c->d_mask;
  if (int rc = solid_mask_device(c, fine, mask, c->st)) return rc;
  HIPCHK(hipEventRecord(c->fork, c->st));
  HIPCHK(hipStreamWaitEvent(c->st2, c->fork, 0));
  silt_sheet_opts o;
  silt_sheet_opts_default(&o);
  o.path = c->path;
  o.row_split = 0;
  o.sub_cap = fine ? c->fine_sub_cap : 8.0f;
  const float nu_w = dynamic_viscosity(3, 1.0f);
  SheetPlanes s = ws->planes();
  s.h = fine ? c->f.water : c->c.water;
  s.u = fine ? c->f.uw : c->c.uw;
  s.v = fine ? c->f.vw : c->c.vw;
  s.bed = fine ? c->f.bed : c->c.bed;
  s.solid = mask;
  if (!fine) core_soften(s.u, s.v, res, c->st);
  displace_occupied(s.h, s.u, s.v, mask, res, s.tg, s.ts, fine ? fine_book(c, 0) : coarse_book(c, 0), c->st);
  silt_drive dw{SILT_DRIVE_GRAVITY, 0.0f, 0.0f, 0.08f};
  int rc = step_device_planes(s, res, cell, dt, dw, o, true, true, nu_w, c->st, fine ? fine_book(c, 0) : coarse_book(c, 0));
  if (rc < 0) return rc;
  SheetPlanes s2 = ws2->planes();
  float* nu = fine ? c->nu_f : c->nu_c;
  core_nu(nu, fine ? c->f.grain : c->c.grain, fine ? c->f.moisture : c->c.moisture, res * res, c->st2);
  s2.h = fine ? c->f.slurry : c->c.slurry;
  s2.u = fine ? c->f.us : c->c.us;
  s2.v = fine ? c->f.vs : c->c.vs;
  s2.bed = fine ? c->f.bed_s : c->c.bed;
  s2.nu = nu;
  s2.solid = mask;
  if (!fine) core_soften(s2.u, s2.v, res, c->st2);
  displace_occupied(s2.h, s2.u, s2.v, mask, res, s2.tg, s2.ts, fine ? fine_book(c, 1) : coarse_book(c, 1), c->st2);
  silt_drive ds{SILT_DRIVE_GRAVITY, 0.0f, 0.0f, 0.05f};
  rc = step_device_planes(s2, res, cell, dt, ds, o, true, false, 0.0f, c->st2, fine ? fine_book(c, 1) : coarse_book(c, 1));
  HIPCHK(hipEventRecord(c->join, c->st2));
  HIPCHK(hipStreamWaitEvent(c->st, c->join, 0));
  return rc < 0 ? rc : 0;
}

// scour_contacts (src/fluid_core.rs:7364-7421) after the window's step.
int scour(silt_core* c, float dt) {
  if (!(dt > 0.0f) || c->solids.empty()) return 0;
  if (int rc = solid_mask_device(c, false, c->d_mask, c->st)) return rc;
  core_scour(c->c, c->d_mask, c->d_bed_drop, dt, c->d_eroded, c->d_hole, source_book(c) + 1, c->st);
  return hipGetLastError() == hipSuccess ? 0 : SILT_EDEVICE;
}

int step_pair(silt_core* c, float dt, bool fine) {
  if (fine && !c->fine_on) return 0;
  if (!c->solids.empty()) return step_pair_walled(c, dt, fine);
  int res = fine ? kFineRes : kWindowRes;
  float cell = fine ? kFineCell : kCoarseCell;
  Workspace* ws = workspace_for(res, 0);
  Workspace* ws2 = workspace_for(res, 1);
  if (!ws || !ws2) return SILT_ENOMEM;
  HIPCHK(hipEventRecord(c->fork, c->st));
  HIPCHK(hipStreamWaitEvent(c->st2, c->fork, 0));
  silt_sheet_opts o;
  silt_sheet_opts_default(&o);
  o.path = c->path;
  o.row_split = 0;
  o.pace = fine ? c->fine_pace : 0;
  o.sub_cap = fine ? c->fine_sub_cap : 8.0f;
  o.confine = fine ? SILT_CONFINE_REFERENCE : c->coarse_confine;
  SheetPlanes s = ws->planes();
  const float nu_w = dynamic_viscosity(3, 1.0f);
  if (!fine) return step_coarse_pair(c, dt, ws, ws2, o, nu_w);
  // Water layer.
  s.h = fine ? c->f.water : c->c.water;
  s.u = fine ? c->f.uw : c->c.uw;
  s.v = fine ? c->f.vw : c->c.vw;
  s.bed = fine ? c->f.bed : c->c.bed;
  silt_drive dw{SILT_DRIVE_GRAVITY, 0.0f, 0.0f, 0.08f};
  // The 2 cm sheet plans from its tiles' tests on the GPU (paces 0 and 1);
  // the whole-wet paces read the planes back.
  auto fine_step = [&](const SheetPlanes& sp, const silt_drive& d, bool nu_uni, float nu0, FineScratch& fsc,
                       hipStream_t q, SheetTotals* book) -> int {
    StepPlan plan;
    int pr = plan_fine_device(sp.h, sp.u, sp.v, sp.bed, res, cell, dt, o.eps, o.sub_cap, o.pace, fsc, &plan, q);
    if (pr < 0) return pr;
    if (pr == 0) return step_device_planes(sp, res, cell, dt, d, o, true, nu_uni, nu0, q, book);
    return plan.boxes.empty() ? 0 : run_plan(sp, plan, res, cell, dt, d, o, true, nu_uni, nu0, q, book);
  };
  int rc = fine_step(s, dw, true, nu_w, c->fs[0], c->st, fine_book(c, 0));
  if (rc < 0) return rc;
  // Slurry layer at its own viscosity, on the second stream.
  SheetPlanes s2 = ws2->planes();
  float* nu = fine ? c->nu_f : c->nu_c;
  core_nu(nu, fine ? c->f.grain : c->c.grain, fine ? c->f.moisture : c->c.moisture, res * res, c->st2);
  if (!fine) core_soften(c->c.us, c->c.vs, res, c->st2);
  s2.h = fine ? c->f.slurry : c->c.slurry;
  s2.u = fine ? c->f.us : c->c.us;
  s2.v = fine ? c->f.vs : c->c.vs;
  s2.bed = fine ? c->f.bed_s : c->c.bed;
  s2.nu = nu;
  silt_drive ds{SILT_DRIVE_GRAVITY, 0.0f, 0.0f, 0.05f};
  rc = fine_step(s2, ds, false, 0.0f, c->fs[1], c->st2, fine_book(c, 1));
  HIPCHK(hipEventRecord(c->join, c->st2));
  HIPCHK(hipStreamWaitEvent(c->st, c->join, 0));
  return rc < 0 ? rc : 0;
}

void peel(silt_core* c, const PushList& pl, bool fine, bool skip_held) {
  GridGeo g{fine ? kFineRes : kWindowRes, fine ? kFineCell : kCoarseCell, fine ? c->fine_origin[0] : c->origin[0],
            fine ? c->fine_origin[1] : c->origin[1]};
  GridGeo cg{kWindowRes, kCoarseCell, c->origin[0], c->origin[1]};
  core_peel(g, fine ? c->f.water : c->c.water, fine ? c->f.uw : c->c.uw, fine ? c->f.vw : c->c.vw, pl, 0, 0, g.res,
            g.res, c->d_hold, cg, skip_held, c->d_ev, c->d_nev, kSprayCap, c->d_acc, fine ? kFineArea : kCoarseArea,
            c->st);
}

void shove(silt_core* c, const std::vector<CorePush>& v, bool fine) {
  if (v.empty()) return;
  GridGeo g{fine ? kFineRes : kWindowRes, fine ? kFineCell : kCoarseCell, fine ? c->fine_origin[0] : c->origin[0],
            fine ? c->fine_origin[1] : c->origin[1]};
  PushList pl = push_list(v);
  if (fine)
    core_push_on(g, c->f.water, c->f.slurry, c->f.uw, c->f.vw, c->f.us, c->f.vs, pl, 0, 0, g.res, g.res, c->st);
  else
    core_push_on(g, c->c.water, c->c.slurry, c->c.uw, c->c.vw, c->c.us, c->c.vs, pl, 0, 0, g.res, g.res, c->st);
}

// The reference's step from the refine to the coarse spray (FluidCore::step,
// src/fluid_core.rs:6034-6128): the 2 cm patch under the tires, then the
// coarse pair.  timed: events 0-1 bracket the patch's own work.
int step_nested(silt_core* c, const std::vector<CorePush>& pushes, const PushList& pl, float dt, bool timed) {
  if (timed) HIPCHK(hipEventRecord(c->tev[0], c->st));
  if (timed) c->timed = 1;
  int rc = refine(c, pushes, dt);
  if (rc) return rc;
  std::vector<CorePush> fine_p, coarse_p;
  for (const CorePush& p : pushes)
    if (cell_is_refined(c, p.x, p.z)) fine_p.push_back(p);
  shove(c, fine_p, true);
  if ((rc = step_pair(c, dt, true))) return rc;
  if (c->fine_on) peel(c, pl, true, false);
  if ((rc = restrict_fine(c))) return rc;
  if (timed) HIPCHK(hipEventRecord(c->tev[1], c->st));
  for (const CorePush& p : pushes)
    if (!fine_held_at(c, p.x, p.z)) coarse_p.push_back(p);
  shove(c, coarse_p, false);
  if ((rc = step_pair(c, dt, false))) return rc;
  peel(c, pl, false, true);
  return scour(c, dt);
}

SurGrid core_grid(const silt_core* c) {
  return SurGrid{c->c.water, c->c.slurry, c->c.uw, c->c.vw, c->c.bed, kWindowRes, kCoarseCell, c->origin[0],
                 c->origin[1]};
}

// Region of the tire sub-grid network for these pushes.
SurRegion sur_region(const silt_core* c, const std::vector<CorePush>& p) {
  return sur_region_at(p.data(), (int)p.size(), c->origin[0], c->origin[1], kCoarseCell, kWindowRes);
}

// Coarse-only step: every tire shoves the coarse sheet, no patch.  With a
// region, the network's fluxes (predicted from the state before the step)
// are applied after it.
int step_coarse(silt_core* c, const std::vector<CorePush>& pushes, const PushList& pl, float dt, bool fade,
                bool surrogate, bool timed) {
  int rc = restrict_fine(c);
  if (rc) return rc;
  if (fade && (rc = near_fade(c, dt))) return rc;
  SurRegion r{0, 0, 0, 0};
  if (surrogate) {
    r = sur_region(c, pushes);
    if (r.nx) {
      if (timed) HIPCHK(hipEventRecord(c->tev[0], c->st));
      if (timed) c->timed = 2;
      sur_context(core_grid(c), r, pl, c->sur_ctx, c->st);
      sur_gather_region(c->sur_ctx, r, c->net.xb, c->st);
      mlp_infer(c->net, (size_t)r.nx * r.nz);
      if (timed) HIPCHK(hipEventRecord(c->tev[1], c->st));
    }
  }
  shove(c, pushes, false);
  if ((rc = step_pair(c, dt, false))) return rc;
  peel(c, pl, false, false);
  if (r.nx) {
    if (timed) HIPCHK(hipEventRecord(c->tev[2], c->st));
    sur_apply(c->c.water, c->c.slurry, kWindowRes, c->net.y, r, c->sur_fac, c->st, coarse_book(c, 0));
    if (timed) HIPCHK(hipEventRecord(c->tev[3], c->st));
  }
  return 0;
}

// Milliseconds of the tire block of the last timed step (after a sync).
float tire_ms(silt_core* c) {
  float a = 0.0f, b = 0.0f;
  if (c->timed >= 1 && hipEventElapsedTime(&a, c->tev[0], c->tev[1]) != hipSuccess) a = 0.0f;
  if (c->timed == 2 && hipEventElapsedTime(&b, c->tev[2], c->tev[3]) != hipSuccess) b = 0.0f;
  return a + b;
}

float* const* core_planes6(silt_core* c, float** out) {
  out[0] = c->c.water;
  out[1] = c->c.slurry;
  out[2] = c->c.uw;
  out[3] = c->c.vw;
  out[4] = c->c.us;
  out[5] = c->c.vs;
  return out;
}

int sur_alloc(silt_core* c) {
  if (c->sur_ctx) return 0;
  bool ok = dalloc(c, &c->sur_ctx, kSurCtxFloats) && dalloc(c, &c->sur_fac, 2 * kSurMax * kSurMax) &&
            dalloc(c, &c->d_err, 1) && dalloc(c, &c->d_sq, 1);
  for (int k = 0; k < 6 && ok; k++) ok = dalloc(c, &c->snap[k], WN);
  for (int k = 0; k < 2 && ok; k++) ok = dalloc(c, &c->judged_h[k], WN);
  for (int k = 0; k < 4 && ok; k++) ok = hipEventCreate(&c->tev[k]) == hipSuccess;
  return ok ? 0 : SILT_ENOMEM;
}

// A surrogate step, judged against the nested one every judge_every steps:
// the surrogate runs from a snapshot, its depths are kept aside, the nested
// step runs from the same snapshot and stays; the largest depth difference
// is the surrogate's error.  Over the tolerance, the next judge_every steps
// are nested only.
int step_surrogate(silt_core* c, const std::vector<CorePush>& pushes, const PushList& pl, float dt) {
  if (c->fallback_left > 0) {
    // After a fallback the next surrogate step is judged again.
    if (--c->fallback_left == 0) c->sur_steps = 0;
    c->last_mode = SILT_TIRE_NESTED;
    return step_nested(c, pushes, pl, dt, true);
  }
  const bool judge = c->judge_every > 0 && c->sur_steps % c->judge_every == 0 && !pushes.empty();
  c->sur_steps++;
  if (!judge) {
    c->last_mode = SILT_TIRE_SURROGATE;
    int rc = step_coarse(c, pushes, pl, dt, true, true, true);
    return rc ? rc : scour(c, dt);
  }
  float* p6[6];
  core_planes6(c, p6);
  int rc = restrict_fine(c);
  if (rc) return rc;
  for (int k = 0; k < 6; k++)
    HIPCHK(hipMemcpyAsync(c->snap[k], p6[k], WN * sizeof(float), hipMemcpyDeviceToDevice, c->st));
  // The judged step's ledger lines are undone with its planes.
  HIPCHK(hipMemcpyAsync(c->d_book + kBookDoubles, c->d_book, sizeof(CoreBook), hipMemcpyDeviceToDevice, c->st));
  if ((rc = step_coarse(c, pushes, pl, dt, false, true, true))) return rc;
  HIPCHK(hipEventSynchronize(c->tev[3]));
  const float surrogate_ms = tire_ms(c);
  c->timed = 0;
  HIPCHK(hipMemcpyAsync(c->judged_h[0], c->c.water, WN * sizeof(float), hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->judged_h[1], c->c.slurry, WN * sizeof(float), hipMemcpyDeviceToDevice, c->st));
  for (int k = 0; k < 6; k++)
    HIPCHK(hipMemcpyAsync(p6[k], c->snap[k], WN * sizeof(float), hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->d_book, c->d_book + kBookDoubles, sizeof(CoreBook), hipMemcpyDeviceToDevice, c->st));
  // The surrogate's spray is undone with its planes (the coarse spray of a
  // step starts from zero; frozen volumes only move in a restrict).
  HIPCHK(hipMemsetAsync(c->d_nev, 0, sizeof(unsigned), c->st));
  HIPCHK(hipMemsetAsync(c->d_acc, 0, sizeof(double), c->st));
  if ((rc = step_nested(c, pushes, pl, dt, false))) return rc;
  HIPCHK(hipMemsetAsync(c->d_err, 0, sizeof(unsigned), c->st));
  HIPCHK(hipMemsetAsync(c->d_sq, 0, sizeof(double), c->st));
  sur_max_abs_diff(c->judged_h[0], c->c.water, WN, c->d_err, c->st);
  sur_max_abs_diff(c->judged_h[1], c->c.slurry, WN, c->d_err, c->st);
  const SurRegion r = sur_region(c, pushes);
  if (r.nx) {
    sur_region_sq_diff(c->judged_h[0], c->c.water, kWindowRes, r, c->d_sq, c->st);
    sur_region_sq_diff(c->judged_h[1], c->c.slurry, kWindowRes, r, c->d_sq, c->st);
  }
  unsigned bits = 0;
  double sq = 0.0;
  HIPCHK(hipMemcpyAsync(&bits, c->d_err, sizeof bits, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(&sq, c->d_sq, sizeof sq, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  memcpy(&c->last_judge_err, &bits, 4);
  c->last_judge_rms = r.nx ? (float)sqrt(sq / (2.0 * r.nx * r.nz)) : 0.0f;
  c->judged++;
  if (!(c->last_judge_err <= c->judge_tol)) {
    c->fallbacks++;
    c->fallback_left = c->judge_every;
  }
  c->last_mode = SILT_TIRE_NESTED;
  c->last_tire_ms = surrogate_ms;
  return 0;
}

// cylinder_segment_volume (src/fluid_core.rs:4606-4615).
float cylinder_segment_volume(float depth, float radius, float width) {
  const float d = clampf(depth, 0.0f, radius * 2.0f);
  if (d <= 1.0e-4f || width <= 0.0f || radius <= 1.0e-3f) return 0.0f;
  const float cc = clampf((radius - d) / radius, -1.0f, 1.0f);
  const float area = radius * radius * acosf(cc) - (radius - d) * sqrtf(fmaxf(2.0f * radius * d - d * d, 0.0f));
  return fmaxf(area, 0.0f) * width;
}

bool solid_contains_host(const silt_hydro_solid& s, float x, float z) {
  const float dx = x - s.x, dz = z - s.z;
  const float fwd_x = sinf(s.yaw), fwd_z = cosf(s.yaw);
  const float right_x = fwd_z, right_z = -fwd_x;
  const float along = dx * fwd_x + dz * fwd_z;
  const float lat = dx * right_x + dz * right_z;
  return fabsf(along) <= fmaxf(s.half_length, 0.0f) && fabsf(lat) <= fmaxf(s.half_width, 0.0f);
}

// hydro_load (src/fluid_core.rs:7164-7249) from host copies of the window's
// water, velocity and bed and its wall mask: the hydrostatic and dynamic
// push on the wall faces, and each solid's displaced water from the mean
// surface of the wet ring around it (ring_surface, :7251-7286).
silt_hydro_load hydro_load_host(const silt_core* c, const std::vector<uint8_t>& mask, const std::vector<float>& water,
                                const std::vector<float>& uw, const std::vector<float>& vw,
                                const std::vector<float>& bedp) {
  silt_hydro_load L{};
  const float cell = kCoarseCell;
  const int res = kWindowRes;
  for (uint8_t m : mask) {
    if (m == 1) L.wheel_blocks++;
    else if (m == 2) L.belly_blocks++;
  }
  static const int nb[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
  float fx = 0.0f, fz = 0.0f;
  for (int j = 0; j < res; j++)
    for (int i = 0; i < res; i++) {
      const size_t idx = (size_t)j * res + i;
      if (mask[idx] != 0) continue;
      const float h = water[idx];
      if (h < 1.0e-4f || !isfinite(h)) continue;
      for (int q = 0; q < 4; q++) {
        const int ni = i + nb[q][0], nj = j + nb[q][1];
        if (ni < 0 || nj < 0 || ni >= res || nj >= res) continue;
        if (mask[(size_t)nj * res + ni] == 0) continue;
        const float nx = (float)nb[q][0], nz = (float)nb[q][1];
        const float hydro = 1000.0f * kGravity * h * h * 0.5f * cell;
        fx += nx * hydro;
        fz += nz * hydro;
        const float vn = uw[idx] * nx + vw[idx] * nz;
        if (vn > 0.0f) {
          const float dyn_f = 0.5f * 1000.0f * vn * vn * h * cell;
          fx += nx * dyn_f;
          fz += nz * dyn_f;
        }
      }
    }
  float wheel_b = 0.0f, hull_b = 0.0f, displaced = 0.0f;
  for (const silt_hydro_solid& s : c->solids) {
    const float reach = fmaxf(s.half_length, s.half_width) + 0.55f;
    const int i0 = (int)floorf((s.x - reach - c->origin[0]) / cell);
    const int i1 = (int)floorf((s.x + reach - c->origin[0]) / cell);
    const int j0 = (int)floorf((s.z - reach - c->origin[1]) / cell);
    const int j1 = (int)floorf((s.z + reach - c->origin[1]) / cell);
    float bed = 0.0f, wat = 0.0f, n = 0.0f;
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) {
        if (i < 0 || j < 0 || i >= res || j >= res) continue;
        float x, z;
        coarse_xz(c, i, j, &x, &z);
        if (solid_contains_host(s, x, z)) continue;
        if (hypotf(x - s.x, z - s.z) > reach) continue;
        const size_t idx = (size_t)j * res + i;
        if (water[idx] < 0.01f) continue;
        bed += bedp[idx];
        wat += water[idx];
        n += 1.0f;
      }
    if (n < 1.0f) continue;
    bed = bed / n;
    wat = wat / n;
    if (s.radius > 0.0f) {
      const float vol = cylinder_segment_volume(wat, s.radius, s.width);
      displaced += vol;
      wheel_b += 1000.0f * kGravity * vol;
    } else {
      const float eta = bed + wat;
      const float height = fmaxf(s.top_y - s.bottom_y, 0.0f);
      const float sub = clampf(eta - s.bottom_y, 0.0f, height);
      const float vol = (s.half_length * 2.0f) * (s.half_width * 2.0f) * sub * 0.36f;
      displaced += vol;
      hull_b += 1000.0f * kGravity * vol;
    }
  }
  L.force_x = fx;
  L.force_z = fz;
  L.buoyancy_n = wheel_b + hull_b;
  L.hull_buoyancy_n = hull_b;
  L.wheel_buoyancy_n = wheel_b;
  L.displaced_m3 = displaced;
  return L;
}

// A tire wakes the tracers at 0.25 m/s or slipping 5 % (step_tracers
// :6215-6219), and its 8 m reach joins the visit box (tracer_visit_box
// :6313-6330).
bool tracer_wake(const std::vector<CorePush>& p) {
  for (const CorePush& q : p)
    if ((isfinite(q.speed) && q.speed >= 0.25f) || (isfinite(q.slip) && fabsf(q.slip) >= 0.05f)) return true;
  return false;
}

SpanBox tracer_push_box(const silt_core* c, const std::vector<CorePush>& p) {
  SpanBox b{kWindowRes, -1, kWindowRes, -1};
  const int reach = (int)ceilf(8.0f / kCoarseCell);
  for (const CorePush& q : p) {
    const int ci = coarse_axis(c->origin[0], q.x), cj = coarse_axis(c->origin[1], q.z);
    if (ci < 0 || cj < 0) continue;
    b.i0 = std::min(b.i0, std::max(ci - reach, 0));
    b.i1 = std::max(b.i1, std::min(ci + reach, kWindowRes - 1));
    b.j0 = std::min(b.j0, std::max(cj - reach, 0));
    b.j1 = std::max(b.j1, std::min(cj + reach, kWindowRes - 1));
  }
  return b;
}

// step_tracers after the window's step (FluidCore::step :6079).
int tracer_pass(silt_core* c, const std::vector<CorePush>& pushes, float dt) {
  if (!(dt > 0.0f) || !isfinite(dt)) return 0;
  TracerPushes tp{};
  tp.n = (int)std::min(pushes.size(), (size_t)kMaxPush);
  for (int k = 0; k < tp.n; k++) tp.p[k] = pushes[k];
  tracer_step(c->tr, c->c, c->d_tstate, tp, tracer_wake(pushes) ? 1 : 0, tracer_push_box(c, pushes), dt,
              expf(-dt / 1.8f), expf(-dt / 4.0f), c->origin[0], c->origin[1], c->st);
  return hipGetLastError() == hipSuccess ? 0 : SILT_EDEVICE;
}

// The ledger's volumes (FluidCore::ledger, src/fluid_core.rs:6821-6883) in
// FP64 over the window and, while it is on, the 2 cm patch; with book, the
// booked lines too, read in the same wait.
int read_ledger(silt_core* c, silt_fluid_ledger* out, CoreBook* book) {
  SheetStats* s = c->d_lstats;
  stats_compute(c->c.water, c->c.uw, c->c.vw, kWindowRes, 1.0e-6f, s, c->st);
  stats_compute(c->c.slurry, c->c.us, c->c.vs, kWindowRes, 1.0e-6f, s + 1, c->st);
  if (c->fine_on) {
    stats_compute(c->f.water, c->f.uw, c->f.vw, kFineRes, 1.0e-6f, s + 2, c->st);
    stats_compute(c->f.slurry, c->f.us, c->f.vs, kFineRes, 1.0e-6f, s + 3, c->st);
  }
  SheetStats h[4]{};
  double acc[4];
  HIPCHK(hipMemcpyAsync(h, s, sizeof h, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(acc, c->d_acc, sizeof acc, hipMemcpyDeviceToHost, c->st));
  if (book) HIPCHK(hipMemcpyAsync(book, c->d_book, sizeof(CoreBook), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  out->water_m3 = h[0].vol * kCoarseArea + (c->fine_on ? h[2].vol * kFineArea : 0.0);
  out->slurry_m3 = h[1].vol * kCoarseArea + (c->fine_on ? h[3].vol * kFineArea : 0.0);
  double pw = 0.0, ps = 0.0;
  for (const auto& kv : c->parked) {
    pw += (double)kv.second.water;
    ps += (double)kv.second.slurry;
  }
  out->frozen_water_m3 = acc[1] + pw * kCoarseArea;
  out->frozen_slurry_m3 = acc[2] + ps * kCoarseArea;
  out->spray_m3 = c->spray_m3;
  out->total_m3 = out->water_m3 + out->slurry_m3 + out->frozen_water_m3 + out->frozen_slurry_m3 + out->spray_m3;
  return SILT_OK;
}

// Each layer's total (water with its frozen part and the spray; slurry with
// its frozen part) and what the booked lines say it gained, in m^3.
void layer_sums(const silt_fluid_ledger& l, const CoreBook& b, double tot[2], double lines[2]) {
  tot[0] = l.water_m3 + l.frozen_water_m3 + l.spray_m3;
  tot[1] = l.slurry_m3 + l.frozen_slurry_m3;
  for (int k = 0; k < 2; k++)
    lines[k] = (b.coarse[k].residue - b.coarse[k].cut) * kCoarseArea +
               (b.fine[k].residue - b.fine[k].cut) * kFineArea + b.handoff[k] + b.source[k];
}

// After a caller edit, the base becomes the current totals less the lines
// booked so far, so the gap starts again from zero.
int ensure_base(silt_core* c) {
  if (!c->rebase) return 0;
  silt_fluid_ledger l{};
  CoreBook b{};
  int rc = read_ledger(c, &l, &b);
  if (rc) return rc;
  double tot[2], lines[2];
  layer_sums(l, b, tot, lines);
  c->base[0] = tot[0] - lines[0];
  c->base[1] = tot[1] - lines[1];
  c->rebase = false;
  return 0;
}

int step_by_model(silt_core* c, const std::vector<CorePush>& pushes, const PushList& pl, float dt) {
  c->timed = 0;
  switch (c->tire_model) {
    case SILT_TIRE_COARSE: {
      c->last_mode = SILT_TIRE_COARSE;
      int rc = step_coarse(c, pushes, pl, dt, true, false, false);
      return rc ? rc : scour(c, dt);
    }
    case SILT_TIRE_SURROGATE:
      return step_surrogate(c, pushes, pl, dt);
    default:
      c->last_mode = SILT_TIRE_NESTED;
      return step_nested(c, pushes, pl, dt, c->tev[0] != nullptr);
  }
}

}  // namespace

extern "C" {

silt_core* silt_core_create(float center_x, float center_z, int path) {
  if (path != SILT_PATH_EXACT && path != SILT_PATH_FAST && path != SILT_PATH_FUSED_EXACT) return nullptr;
  silt_core* c = new silt_core();
  c->path = path;
  c->center[0] = center_x;
  c->center[1] = center_z;
  float extent = (float)kWindowRes * kCoarseCell;
  c->origin[0] = floorf((center_x - extent * 0.5f) / kCoarseCell) * kCoarseCell;
  c->origin[1] = floorf((center_z - extent * 0.5f) / kCoarseCell) * kCoarseCell;
  c->hold.assign(WN, 0);
  bool ok = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&c->join, hipEventDisableTiming) == hipSuccess;
  float** cp[8] = {&c->c.water, &c->c.slurry, &c->c.uw, &c->c.vw, &c->c.us, &c->c.vs, &c->c.bed, &c->c.moisture};
  for (auto p : cp) ok = ok && dalloc(c, p, WN);
  float** fp[9] = {&c->f.water, &c->f.slurry, &c->f.uw, &c->f.vw, &c->f.us, &c->f.vs, &c->f.bed, &c->f.bed_s,
                   &c->f.moisture};
  for (auto p : fp) ok = ok && dalloc(c, p, FN);
  ok = ok && dalloc(c, &c->c.grain, WN) && dalloc(c, &c->f.grain, FN) && dalloc(c, &c->d_hold, WN) &&
       dalloc(c, &c->nu_c, WN) && dalloc(c, &c->nu_f, FN) && near_alloc(c, &c->near[0]) && near_alloc(c, &c->near[1]) &&
       dalloc(c, &c->d_cols, 1024) && dalloc(c, &c->d_ev, kSprayCap) && dalloc(c, &c->d_nev, 1) &&
       dalloc(c, &c->d_acc, 4) && dalloc(c, &c->soil_bed, WN) && dalloc(c, &c->soil_placed, WN) &&
       dalloc(c, &c->soil_init, WN) && dalloc(c, &c->soil_rut, WN) && dalloc(c, &c->d_contacts, kMaxPush) &&
       dalloc(c, &c->d_stats, 2) && dalloc(c, &c->d_lstats, 4) && dalloc(c, &c->d_pub, 13 * WN) &&
       dalloc(c, &c->d_book, 2 * kBookDoubles) && dalloc(c, &c->d_solids, kMaxSolids) &&
       dalloc(c, &c->d_mask, WN) && dalloc(c, &c->d_fmask, FN) && dalloc(c, &c->d_bed_drop, WN) &&
       dalloc(c, &c->d_eroded, 1) && dalloc(c, &c->d_hole, 1) && dalloc(c, &c->tr.foam, WN) &&
       dalloc(c, &c->tr.turbid, WN) && dalloc(c, &c->tr.foam_live, WN) && dalloc(c, &c->tr.wet, WN) &&
       dalloc(c, &c->tr.shore, WN) && dalloc(c, &c->tr.scratch, 2 * WN) && dalloc(c, &c->tr.acc, 2) &&
       dalloc(c, &c->d_tstate, 1) && dalloc(c, &c->pond.run, 2 * WN) &&
       dalloc(c, &c->pond.pre, 2 * (kWindowRes + 1) * (kWindowRes + 1)) && dalloc(c, &c->pond.f0, WN) &&
       dalloc(c, &c->pond.f1, WN) && dalloc(c, &c->pond.f2, WN) && dalloc(c, &c->pond.f3, WN) &&
       dalloc(c, &c->pond.f4, WN) && dalloc(c, &c->pond.old_w, WN) && dalloc(c, &c->pond.old_s, WN) &&
       dalloc(c, &c->pond.m0, WN) && dalloc(c, &c->pond.m1, WN) && dalloc(c, &c->pond.ring, WN) &&
       dalloc(c, &c->d_pond, 1) && dalloc(c, &c->sus.susp, WN) && dalloc(c, &c->sus.give, WN) &&
       dalloc(c, &c->sus.conc, WN) && dalloc(c, &c->sus.accept, WN) && dalloc(c, &c->sus.taken, WN) &&
       dalloc(c, &c->sus.dir, WN) && dalloc(c, &c->sus.acc, 2) &&
       hipHostMalloc((void**)&c->h_stats, 2 * sizeof(SheetStats), hipHostMallocDefault) == hipSuccess;
  if (ok) {
    // Grain::Other everywhere (FluidCore::new, src/fluid_core.rs:4979-4980).
    ok = hipMemset(c->c.grain, 4, WN) == hipSuccess && hipMemset(c->f.grain, 4, FN) == hipSuccess;
    core_near_clear(c->near[0], c->st);
    core_near_clear(c->near[1], c->st);
    ok = ok && hipStreamSynchronize(c->st) == hipSuccess;
  }
  ok = ok && sur_alloc(c) == 0;
  if (!ok) {
    silt_core_destroy(c);
    return nullptr;
  }
  return c;
}

void silt_core_destroy(silt_core* c) {
  if (!c) return;
  if (c->st) (void)hipStreamSynchronize(c->st);
  if (c->net_ready) mlp_free(c->net);
  for (hipEvent_t e : c->tev)
    if (e) (void)hipEventDestroy(e);
  if (c->st2) (void)hipStreamSynchronize(c->st2);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->h_stats) (void)hipHostFree(c->h_stats);
  plan_labeller_free(c->lab);
  for (PlanScratch& q : c->ps) plan_scratch_free(q);
  for (FineScratch& q : c->fs) fine_scratch_free(q);
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->join) (void)hipEventDestroy(c->join);
  if (c->st2) (void)hipStreamDestroy(c->st2);
  if (c->st) (void)hipStreamDestroy(c->st);
  delete c;
}

int silt_core_set_fine_pace(silt_core* c, int pace) {
  if (!c || pace < 0) return SILT_EINVAL;
  c->fine_pace = pace > 3 ? 3 : pace;
  return SILT_OK;
}

int silt_core_set_fine_sub_cap(silt_core* c, float cap) {
  if (!c) return SILT_EINVAL;
  c->fine_sub_cap = isfinite(cap) ? clampf(cap, 1.0f, 8.0f) : 8.0f;
  return SILT_OK;
}

int silt_core_set_near_field(silt_core* c, int on) {
  if (!c) return SILT_EINVAL;
  c->near_on = on != 0;
  if (!c->near_on) {
    core_near_clear(c->near[c->near_cur], c->st);
    HIPCHK(hipStreamSynchronize(c->st));
  }
  return SILT_OK;
}

int silt_core_set_coarse_confine(silt_core* c, int confine) {
  if (!c || (confine != SILT_CONFINE_REFERENCE && confine != SILT_CONFINE_JACOBI)) return SILT_EINVAL;
  c->coarse_confine = confine;
  return SILT_OK;
}

int silt_core_window_origin(const silt_core* c, float* out2) {
  if (!c || !out2) return SILT_EINVAL;
  out2[0] = c->origin[0];
  out2[1] = c->origin[1];
  return SILT_OK;
}

int silt_core_get_planes(silt_core* c, float* water, float* slurry, float* uw, float* vw, float* us, float* vs,
                         float* bed, float* moisture) {
  if (!c) return SILT_EINVAL;
  float* dst[8] = {water, slurry, uw, vw, us, vs, bed, moisture};
  float* src[8] = {c->c.water, c->c.slurry, c->c.uw, c->c.vw, c->c.us, c->c.vs, c->c.bed, c->c.moisture};
  for (int k = 0; k < 8; k++)
    if (dst[k]) HIPCHK(hipMemcpyAsync(dst[k], src[k], WN * sizeof(float), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_set_planes(silt_core* c, const float* water, const float* slurry, const float* uw, const float* vw,
                         const float* us, const float* vs, const float* bed, const float* moisture) {
  if (!c) return SILT_EINVAL;
  c->rebase = true;
  if (restrict_fine(c)) return SILT_EDEVICE;
  const float* src[8] = {water, slurry, uw, vw, us, vs, bed, moisture};
  float* dst[8] = {c->c.water, c->c.slurry, c->c.uw, c->c.vw, c->c.us, c->c.vs, c->c.bed, c->c.moisture};
  for (int k = 0; k < 8; k++)
    if (src[k]) HIPCHK(hipMemcpyAsync(dst[k], src[k], WN * sizeof(float), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

// The setters below run on host copies of the planes they touch (they are
// not on the per-step path): read, change as the reference does, write back.
int silt_core_map_water_bed(silt_core* c, const float* bed, const float* water) {
  if (!c || !bed || !water) return SILT_EINVAL;
  c->rebase = true;
  if (restrict_fine(c)) return SILT_EDEVICE;
  std::vector<float> b(WN), w(WN), m(WN);
  std::vector<uint8_t> g(WN);
  HIPCHK(hipMemcpyAsync(m.data(), c->c.moisture, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(g.data(), c->c.grain, WN, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int k = 0; k < WN; k++) {
    b[k] = isfinite(bed[k]) ? bed[k] : 0.0f;
    float wv = isfinite(water[k]) ? fmaxf(water[k], 0.0f) : 0.0f;
    w[k] = wv;
    if (wv > 0.0f) {
      m[k] = fmaxf(m[k], 0.8f);
      if (g[k] == 4) g[k] = 1;
    }
  }
  HIPCHK(hipMemcpyAsync(c->c.bed, b.data(), WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->c.water, w.data(), WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->c.moisture, m.data(), WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->c.grain, g.data(), WN, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_pour(silt_core* c, float x0, float x1, float z0, float z1, float depth, int water, int grain,
                   float moisture) {
  if (!c || grain < 0 || grain > 4) return SILT_EINVAL;
  c->rebase = true;
  float d = depth > 0.0f ? depth : 0.0f;
  std::vector<float> h(WN), m(WN);
  std::vector<uint8_t> g(WN);
  float* plane = water ? c->c.water : c->c.slurry;
  HIPCHK(hipMemcpyAsync(h.data(), plane, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(m.data(), c->c.moisture, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(g.data(), c->c.grain, WN, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int j = 0; j < kWindowRes; j++)
    for (int i = 0; i < kWindowRes; i++) {
      float x, z;
      coarse_xz(c, i, j, &x, &z);
      if (x < x0 || x >= x1 || z < z0 || z >= z1) continue;
      int k = j * kWindowRes + i;
      h[k] = d;
      if (!water) {
        g[k] = (uint8_t)grain;
        m[k] = clampf(moisture, 0.0f, 1.0f);
      }
    }
  HIPCHK(hipMemcpyAsync(plane, h.data(), WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->c.moisture, m.data(), WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->c.grain, g.data(), WN, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_lay_slurry(silt_core* c, const float* depth, int grain, float moisture) {
  if (!c || !depth || grain < 0 || grain > 4) return SILT_EINVAL;
  c->rebase = true;
  if (restrict_fine(c)) return SILT_EDEVICE;
  std::vector<float> s(WN), m(WN);
  std::vector<uint8_t> g(WN);
  HIPCHK(hipMemcpyAsync(s.data(), c->c.slurry, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(m.data(), c->c.moisture, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(g.data(), c->c.grain, WN, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  float mm = clampf(moisture, 0.0f, 1.0f);
  for (int k = 0; k < WN; k++) {
    float d = depth[k];
    if (!(isfinite(d) && d > 0.0f)) continue;
    s[k] += d;
    g[k] = (uint8_t)grain;
    m[k] = fmaxf(m[k], mm);
  }
  HIPCHK(hipMemcpyAsync(c->c.slurry, s.data(), WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->c.moisture, m.data(), WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->c.grain, g.data(), WN, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_read_substrate(silt_core* c, const float* bed, const float* moisture, const uint8_t* grain) {
  if (!c || !bed || !moisture || !grain) return SILT_EINVAL;
  c->rebase = true;
  if (restrict_fine(c)) return SILT_EDEVICE;
  std::vector<float> s(WN), m(WN);
  std::vector<uint8_t> g(WN);
  HIPCHK(hipMemcpyAsync(s.data(), c->c.slurry, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(g.data(), c->c.grain, WN, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  // The most common moisture, by 1 % buckets (rounded half away from zero),
  // is the ambient moisture of cells that slide onto the window.
  unsigned hist[101] = {};
  for (int k = 0; k < WN; k++) {
    m[k] = clampf(moisture[k], 0.0f, 1.0f);
    const float pct = m[k] * 100.0f;
    const int whole = (int)pct;
    const int bucket = pct - (float)whole >= 0.5f ? whole + 1 : whole;
    hist[bucket < 100 ? bucket : 100]++;
    if (s[k] > 1.0e-5f) g[k] = grain[k];
  }
  int best = 0;
  unsigned best_n = 0;
  for (int k = 0; k <= 100; k++)
    if (hist[k] > best_n) {
      best = k;
      best_n = hist[k];
    }
  c->moisture_ambient = (float)best * 0.01f;
  HIPCHK(hipMemcpyAsync(c->c.bed, bed, WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->c.moisture, m.data(), WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->c.grain, g.data(), WN, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_push_tire(silt_core* c, const silt_tire_push* p) {
  if (!c || !p) return SILT_EINVAL;
  if (c->pushes.size() >= (size_t)kMaxPush) return SILT_EINVAL;
  c->pushes.push_back(CorePush{p->x, p->z, p->forward_x, p->forward_z, p->radius, p->width, p->speed, p->slip});
  return SILT_OK;
}

int silt_core_step(silt_core* c, float dt, silt_spray_event* events, size_t cap, size_t* nevents) {
  if (!c) return SILT_EINVAL;
  if (int rb = ensure_base(c)) return rb;
  std::vector<CorePush> pushes;
  pushes.swap(c->pushes);
  PushList pl = push_list(pushes);
  hipStream_t st = c->st;
  HIPCHK(hipMemsetAsync(c->d_nev, 0, sizeof(unsigned), st));
  HIPCHK(hipMemsetAsync(c->d_acc, 0, sizeof(double), st));
  const bool live_dt = dt > 0.0f && isfinite(dt);
  // Lift from the column the tire rolls into, before the sheet shoves it
  // (entrain_wake, FluidCore::step :6050-6053).
  if (live_dt && !pushes.empty()) {
    TracerPushes tp{};
    tp.n = (int)std::min(pushes.size(), (size_t)kMaxPush);
    for (int k = 0; k < tp.n; k++) tp.p[k] = pushes[k];
    susp_entrain(c->sus, c->c.water, tp, dt, c->origin[0], c->origin[1], st);
  }
  int rc = step_by_model(c, pushes, pl, dt);
  if (rc) return rc;
  if ((rc = tracer_pass(c, pushes, dt))) return rc;
  // step_suspended (:6086-6089, :8750-8756)
  if (live_dt) susp_step(c->sus, c->c.water, c->c.vw, dt, st);
  c->last_pushes = pushes;
  unsigned nev = 0;
  double vol = 0.0;
  HIPCHK(hipMemcpyAsync(&nev, c->d_nev, sizeof nev, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&vol, c->d_acc, sizeof vol, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (c->timed) c->last_tire_ms = tire_ms(c);
  c->spray_m3 += vol;
  size_t take = std::min((size_t)std::min(nev, kSprayCap), cap);
  if (events && take) {
    std::vector<SprayEvent> ev(take);
    HIPCHK(hipMemcpy(ev.data(), c->d_ev, take * sizeof(SprayEvent), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < take; k++)
      events[k] = silt_spray_event{ev[k].x, ev[k].y, ev[k].z, ev[k].vx, ev[k].vy, ev[k].vz, ev[k].volume, ev[k].kind};
  }
  if (nevents) *nevents = nev;
  // Tombstones of faded tiles pile up in the open-addressing table: move the
  // live tiles into the spare table now and then.
  if (++c->steps % 256 == 0) {
    core_near_rehash(c->near[c->near_cur], c->near[1 - c->near_cur], st);
    c->near_cur = 1 - c->near_cur;
  }
  if (hipGetLastError() != hipSuccess) return SILT_EDEVICE;
  return SILT_OK;
}

int silt_core_ledger(silt_core* c, silt_fluid_ledger* out) {
  if (!c || !out) return SILT_EINVAL;
  return read_ledger(c, out, nullptr);
}

int silt_core_ledger_lines(silt_core* c, silt_core_lines* out) {
  if (!c || !out) return SILT_EINVAL;
  int rc = ensure_base(c);
  if (rc) return rc;
  silt_fluid_ledger l{};
  CoreBook b{};
  if ((rc = read_ledger(c, &l, &b))) return rc;
  double tot[2], lines[2];
  layer_sums(l, b, tot, lines);
  for (int k = 0; k < 2; k++) {
    out->base_m3[k] = c->base[k];
    out->cut_m3[k] = b.coarse[k].cut * kCoarseArea + b.fine[k].cut * kFineArea;
    out->residue_m3[k] = b.coarse[k].residue * kCoarseArea + b.fine[k].residue * kFineArea;
    out->handoff_m3[k] = b.handoff[k];
    out->source_m3[k] = b.source[k];
    out->total_m3[k] = tot[k];
    out->gap_m3[k] = tot[k] - c->base[k] - lines[k];
  }
  return SILT_OK;
}

int silt_core_set_ground(silt_core* c, const float* bed, const float* placed) {
  if (!c || !bed) return SILT_EINVAL;
  if (restrict_fine(c)) return SILT_EDEVICE;
  HIPCHK(hipMemcpyAsync(c->soil_bed, bed, WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->soil_init, bed, WN * 4, hipMemcpyHostToDevice, c->st));
  if (placed)
    HIPCHK(hipMemcpyAsync(c->soil_placed, placed, WN * 4, hipMemcpyHostToDevice, c->st));
  else
    core_fill(c->soil_placed, 0.0f, WN, c->st);
  core_ground(c->soil_bed, c->soil_placed, c->soil_init, c->c.bed, c->soil_rut, WN, c->st);
  c->ground = true;
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_stamp_ruts(silt_core* c, const silt_tire_contact* contacts, size_t n, float dt) {
  if (!c || (!contacts && n)) return SILT_EINVAL;
  if (!c->ground) return SILT_EINVAL;
  if (n > (size_t)kMaxPush) return SILT_EINVAL;
  if (n == 0 || !(dt > 0.0f)) return SILT_OK;
  if (restrict_fine(c)) return SILT_EDEVICE;
  int rc = rut_stamp(c->soil_bed, c->soil_placed, c->soil_init, kWindowRes, kCoarseCell, c->origin[0], c->origin[1],
                     contacts, n, dt, c->d_contacts, c->st);
  if (rc) return rc;
  core_ground(c->soil_bed, c->soil_placed, c->soil_init, c->c.bed, c->soil_rut, WN, c->st);
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_get_ground(silt_core* c, float* bed, float* placed, float* rut) {
  if (!c) return SILT_EINVAL;
  if (!c->ground) return SILT_EINVAL;
  if (bed) HIPCHK(hipMemcpyAsync(bed, c->soil_bed, WN * 4, hipMemcpyDeviceToHost, c->st));
  if (placed) HIPCHK(hipMemcpyAsync(placed, c->soil_placed, WN * 4, hipMemcpyDeviceToHost, c->st));
  if (rut) HIPCHK(hipMemcpyAsync(rut, c->soil_rut, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_publish(silt_core* c, const float* soil_bed, const float* soil_placed, const float* soil_rut,
                      float* ground_rut4, float* film2, float* flow4) {
  if (!c) return SILT_EINVAL;
  float* d_out = c->d_pub;
  float* d_soil = c->d_pub + 10 * WN;
  const float* dsoil[3] = {nullptr, nullptr, nullptr};
  if (c->ground && !soil_bed && !soil_placed && !soil_rut) {
    // The core's own ground: published without a round trip through the host.
    dsoil[0] = c->soil_bed;
    dsoil[1] = c->soil_placed;
    dsoil[2] = c->soil_rut;
  } else {
    const float* soils[3] = {soil_bed, soil_placed, soil_rut};
    for (int k = 0; k < 3; k++)
      if (soils[k]) {
        HIPCHK(hipMemcpyAsync(d_soil + (size_t)k * WN, soils[k], WN * 4, hipMemcpyHostToDevice, c->st));
        dsoil[k] = d_soil + (size_t)k * WN;
      }
  }
  core_publish(c->c, dsoil[0], dsoil[1], dsoil[2], c->sus.susp, c->tr.turbid, d_out, d_out + 4 * WN, d_out + 6 * WN,
               c->st);
  if (ground_rut4) HIPCHK(hipMemcpyAsync(ground_rut4, d_out, 4 * WN * 4, hipMemcpyDeviceToHost, c->st));
  if (film2) HIPCHK(hipMemcpyAsync(film2, d_out + 4 * WN, 2 * WN * 4, hipMemcpyDeviceToHost, c->st));
  if (flow4) HIPCHK(hipMemcpyAsync(flow4, d_out + 6 * WN, 4 * WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_publish_surface(silt_core* c, const float* soil_wetness, const float* soil_placed, float* moisture1,
                              float* wet_edge1, float* foam_crest2) {
  if (!c) return SILT_EINVAL;
  // Scratch: the two optional soil inputs, then the three outputs.
  float* d = c->d_pub;
  const float *d_wet = nullptr, *d_placed = nullptr;
  if (soil_wetness) {
    HIPCHK(hipMemcpyAsync(d, soil_wetness, WN * 4, hipMemcpyHostToDevice, c->st));
    d_wet = d;
  }
  if (soil_placed) {
    HIPCHK(hipMemcpyAsync(d + WN, soil_placed, WN * 4, hipMemcpyHostToDevice, c->st));
    d_placed = d + WN;
  } else if (c->ground) {
    d_placed = c->soil_placed;  // the core's own ground, as silt_core_publish publishes it
  }
  core_publish_surface(c->c, d_wet, d_placed, c->tr.foam, kWindowRes, d + 2 * WN, d + 3 * WN, d + 4 * WN, c->st);
  if (moisture1) HIPCHK(hipMemcpyAsync(moisture1, d + 2 * WN, WN * 4, hipMemcpyDeviceToHost, c->st));
  if (wet_edge1) HIPCHK(hipMemcpyAsync(wet_edge1, d + 3 * WN, WN * 4, hipMemcpyDeviceToHost, c->st));
  if (foam_crest2) HIPCHK(hipMemcpyAsync(foam_crest2, d + 4 * WN, 2 * WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

namespace {

// Orders work on the caller's stream with the handle's: the handle's stream
// waits for what the caller enqueued so far (device inputs), and (after) the
// caller's stream waits for the handle's (device outputs).
int fence_from(silt_core* c, hipStream_t user) {
  HIPCHK(hipEventRecord(c->join, user));
  HIPCHK(hipStreamWaitEvent(c->st, c->join, 0));
  return 0;
}
int fence_to(silt_core* c, hipStream_t user) {
  HIPCHK(hipEventRecord(c->fork, c->st));
  HIPCHK(hipStreamWaitEvent(user, c->fork, 0));
  return 0;
}

}  // namespace

int silt_core_publish_device(silt_core* c, const float* soil_bed, const float* soil_placed, const float* soil_rut,
                             float* ground_rut4, float* film2, float* flow4, void* hip_stream) {
  if (!c) return SILT_EINVAL;
  hipStream_t user = (hipStream_t)hip_stream;
  if (fence_from(c, user)) return SILT_EDEVICE;
  const float* dsoil[3] = {soil_bed, soil_placed, soil_rut};
  if (c->ground && !soil_bed && !soil_placed && !soil_rut) {
    dsoil[0] = c->soil_bed;
    dsoil[1] = c->soil_placed;
    dsoil[2] = c->soil_rut;
  }
  // Outputs the caller does not want land in the handle's scratch.
  float* d_out = c->d_pub;
  core_publish(c->c, dsoil[0], dsoil[1], dsoil[2], c->sus.susp, c->tr.turbid, ground_rut4 ? ground_rut4 : d_out,
               film2 ? film2 : d_out + 4 * WN, flow4 ? flow4 : d_out + 6 * WN, c->st);
  if (hipGetLastError() != hipSuccess) return SILT_EDEVICE;
  return fence_to(c, user) ? SILT_EDEVICE : SILT_OK;
}

int silt_core_publish_surface_device(silt_core* c, const float* soil_wetness, const float* soil_placed,
                                     float* moisture1, float* wet_edge1, float* foam_crest2, void* hip_stream) {
  if (!c) return SILT_EINVAL;
  hipStream_t user = (hipStream_t)hip_stream;
  if (fence_from(c, user)) return SILT_EDEVICE;
  const float* d_placed = soil_placed ? soil_placed : (c->ground ? c->soil_placed : nullptr);
  float* d = c->d_pub;
  core_publish_surface(c->c, soil_wetness, d_placed, c->tr.foam, kWindowRes, moisture1 ? moisture1 : d + 2 * WN,
                       wet_edge1 ? wet_edge1 : d + 3 * WN, foam_crest2 ? foam_crest2 : d + 4 * WN, c->st);
  if (hipGetLastError() != hipSuccess) return SILT_EDEVICE;
  return fence_to(c, user) ? SILT_EDEVICE : SILT_OK;
}

int silt_core_fine_water_debug(silt_core* c, const silt_tire_push* pushes, size_t n, float* water, float* uw, float* vw,
                               float* origin2) {
  if (!c || (!pushes && n) || !water || !origin2 || n > (size_t)kMaxPush) return SILT_EINVAL;
  std::vector<CorePush> p;
  for (size_t k = 0; k < n; k++)
    p.push_back(CorePush{pushes[k].x, pushes[k].z, pushes[k].forward_x, pushes[k].forward_z, pushes[k].radius,
                         pushes[k].width, pushes[k].speed, pushes[k].slip});
  int rc = refine(c, p, 0.0f);
  if (rc) return rc;
  size_t fb = (size_t)FN * sizeof(float);
  HIPCHK(hipMemcpyAsync(water, c->f.water, fb, hipMemcpyDeviceToHost, c->st));
  if (uw) HIPCHK(hipMemcpyAsync(uw, c->f.uw, fb, hipMemcpyDeviceToHost, c->st));
  if (vw) HIPCHK(hipMemcpyAsync(vw, c->f.vw, fb, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  origin2[0] = c->fine_origin[0];
  origin2[1] = c->fine_origin[1];
  return restrict_fine(c);
}

long silt_core_near_tiles(silt_core* c) {
  if (!c) return SILT_EINVAL;
  unsigned n = 0;
  if (hipMemcpyAsync(&n, c->near[c->near_cur].count, sizeof n, hipMemcpyDeviceToHost, c->st) != hipSuccess)
    return SILT_EDEVICE;
  if (hipStreamSynchronize(c->st) != hipSuccess) return SILT_EDEVICE;
  return (long)n;
}

int silt_core_set_tire_model(silt_core* c, int model) {
  if (!c || model < SILT_TIRE_NESTED || model > SILT_TIRE_SURROGATE) return SILT_EINVAL;
  if (model == SILT_TIRE_SURROGATE && !c->net_ready) return SILT_EINVAL;
  c->tire_model = model;
  c->fallback_left = 0;
  c->sur_steps = 0;
  return SILT_OK;
}

int silt_core_set_surrogate(silt_core* c, const float* params, int judge_every, float tol) {
  if (!c || !params || judge_every < 0 || !(tol >= 0.0f)) return SILT_EINVAL;
  if (!c->net_ready) {
    if (mlp_init(c->net, (size_t)kSurMax * kSurMax, 0u, c->st) != 0) return SILT_ENOMEM;
    c->net_ready = true;
  }
  HIPCHK(hipMemcpyAsync(c->net.master, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, c->st));
  mlp_sync_bf16(c->net);
  HIPCHK(hipStreamSynchronize(c->st));
  c->judge_every = judge_every;
  c->judge_tol = tol;
  c->fallback_left = 0;
  c->sur_steps = 0;
  return SILT_OK;
}

int silt_core_surrogate_stats(const silt_core* c, silt_surrogate_stats* out) {
  if (!c || !out) return SILT_EINVAL;
  out->last_tire_ms = c->last_tire_ms;
  out->last_judge_err_m = c->last_judge_err;
  out->last_judge_rms_m = c->last_judge_rms;
  out->judged = c->judged;
  out->fallbacks = c->fallbacks;
  out->last_mode = c->last_mode;
  return SILT_OK;
}

int silt_core_map_water_velocity(silt_core* c, const float* u, const float* v) {
  if (!c || !u || !v) return SILT_EINVAL;
  if (restrict_fine(c)) return SILT_EDEVICE;
  std::vector<float> w(WN), uw(WN), vw(WN);
  HIPCHK(hipMemcpyAsync(w.data(), c->c.water, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(uw.data(), c->c.uw, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(vw.data(), c->c.vw, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int k = 0; k < WN; k++) {
    if (w[k] <= 1.0e-6f) continue;
    uw[k] = isfinite(u[k]) ? u[k] : 0.0f;
    vw[k] = isfinite(v[k]) ? v[k] : 0.0f;
  }
  HIPCHK(hipMemcpyAsync(c->c.uw, uw.data(), WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->c.vw, vw.data(), WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

long silt_core_non_finite_count(silt_core* c) {
  if (!c) return SILT_EINVAL;
  if (restrict_fine(c)) return SILT_EDEVICE;
  float* planes[8] = {c->c.water, c->c.slurry, c->c.uw, c->c.vw, c->c.us, c->c.vs, c->c.bed, c->c.moisture};
  std::vector<float> h(WN);
  long n = 0;
  for (float* p : planes) {
    HIPCHK(hipMemcpyAsync(h.data(), p, WN * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    for (float x : h) n += isfinite(x) ? 0 : 1;
  }
  return n;
}

int silt_core_set_center(silt_core* c, float x, float z) {
  if (!c) return SILT_EINVAL;
  c->center[0] = x;
  c->center[1] = z;
  return SILT_OK;
}

int silt_core_center(const silt_core* c, float* out2) {
  if (!c || !out2) return SILT_EINVAL;
  out2[0] = c->center[0];
  out2[1] = c->center[1];
  return SILT_OK;
}

int silt_core_coarse_cell(silt_core* c, size_t i, size_t j, silt_coarse_cell* out) {
  if (!c || !out) return SILT_EINVAL;
  if (i >= (size_t)kWindowRes || j >= (size_t)kWindowRes) return 0;
  if (restrict_fine(c)) return SILT_EDEVICE;
  const size_t idx = j * kWindowRes + i;
  float v[7];
  float* planes[7] = {c->c.water, c->c.slurry, c->c.uw, c->c.us, c->c.vw, c->c.vs, c->c.bed};
  for (int k = 0; k < 7; k++) HIPCHK(hipMemcpyAsync(&v[k], planes[k] + idx, 4, hipMemcpyDeviceToHost, c->st));
  float m = 0.0f, load[2] = {0.0f, 0.0f};
  HIPCHK(hipMemcpyAsync(&m, c->c.moisture + idx, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(&load[0], c->sus.susp + idx, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(&load[1], c->tr.turbid + idx, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  auto fin = [](float x) { return isfinite(x) ? x : 0.0f; };
  const float wm = fin(v[0]), sm = fin(v[1]), film = wm + sm;
  out->x = c->origin[0] + ((float)i + 0.5f) * kCoarseCell;
  out->z = c->origin[1] + ((float)j + 0.5f) * kCoarseCell;
  out->bed_m = fin(v[6]);
  out->water_m = wm;
  out->slurry_m = sm;
  out->moisture = fin(m);
  out->u = film < 1.0e-5f ? 0.0f : fin((v[2] * wm + v[3] * sm) / film);
  out->v = film < 1.0e-5f ? 0.0f : fin((v[4] * wm + v[5] * sm) / film);
  out->settled_dust_m = 0.0f;
  // wheel-lifted plus stirred solids over the water (:7025-7032)
  float turb = 0.0f;
  if (!(wm < 1.0e-5f)) {
    turb = (fmaxf(fin(load[0]), 0.0f) + fmaxf(fin(load[1]), 0.0f)) / wm;
    turb = turb < 0.0f ? 0.0f : (turb > 1.0f ? 1.0f : turb);
  }
  out->turbidity = turb;
  return 1;
}

int silt_core_sample(silt_core* c, float x, float z, silt_ground_sample* out) {
  if (!c || !out) return SILT_EINVAL;
  *out = silt_ground_sample{};
  if (restrict_fine(c)) return SILT_EDEVICE;
  const int i = coarse_axis(c->origin[0], x), j = coarse_axis(c->origin[1], z);
  if (i < 0 || j < 0) return SILT_OK;
  const size_t idx = (size_t)j * kWindowRes + i;
  float* planes[6] = {c->c.water, c->c.slurry, c->c.us, c->c.vs, c->c.moisture, c->c.bed};
  float v[6];
  for (int k = 0; k < 6; k++) HIPCHK(hipMemcpyAsync(&v[k], planes[k] + idx, 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  out->water_m = v[0];
  out->slurry_m = v[1];
  out->u = v[2];
  out->v = v[3];
  out->settled_dust_m = 0.0f;
  out->moisture = v[4];
  out->bed_m = v[5];
  return SILT_OK;
}

int silt_core_set_hydro_solids(silt_core* c, const silt_hydro_solid* solids, size_t n) {
  if (!c || (!solids && n)) return SILT_EINVAL;
  if (n > (size_t)kMaxSolids) return SILT_EINVAL;
  c->solids.assign(solids, solids + n);
  return SILT_OK;
}

int silt_core_solid_mask(silt_core* c, int fine, uint8_t* out) {
  if (!c || !out) return SILT_EINVAL;
  if (fine && !c->fine_on) return SILT_EINVAL;
  uint8_t* mask = fine ? c->d_fmask : c->d_mask;
  if (solid_mask_device(c, fine != 0, mask, c->st)) return SILT_EDEVICE;
  HIPCHK(hipMemcpyAsync(out, mask, fine ? FN : WN, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_hydro_load(silt_core* c, silt_hydro_load* out) {
  if (!c || !out) return SILT_EINVAL;
  std::vector<uint8_t> mask(WN, 0);
  std::vector<float> w(WN), uw(WN), vw(WN), bed(WN);
  if (!c->solids.empty()) {
    if (solid_mask_device(c, false, c->d_mask, c->st)) return SILT_EDEVICE;
    HIPCHK(hipMemcpyAsync(mask.data(), c->d_mask, WN, hipMemcpyDeviceToHost, c->st));
  }
  HIPCHK(hipMemcpyAsync(w.data(), c->c.water, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(uw.data(), c->c.uw, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(vw.data(), c->c.vw, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(bed.data(), c->c.bed, WN * 4, hipMemcpyDeviceToHost, c->st));
  double eroded = 0.0;
  unsigned hole = 0;
  HIPCHK(hipMemcpyAsync(&eroded, c->d_eroded, sizeof eroded, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(&hole, c->d_hole, sizeof hole, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  *out = hydro_load_host(c, mask, w, uw, vw, bed);
  out->eroded_m3 = (float)eroded;
  memcpy(&out->hole_m, &hole, 4);
  return SILT_OK;
}

int silt_core_note_surface_foam(silt_core* c, float dt) {
  if (!c) return SILT_EINVAL;
  if (!(dt > 0.0f) || !isfinite(dt)) return SILT_OK;
  tracer_note_foam(c->tr, c->c, c->d_tstate, tracer_wake(c->last_pushes) ? 1 : 0, tracer_push_box(c, c->last_pushes),
                   dt, c->st);
  HIPCHK(hipStreamSynchronize(c->st));
  return hipGetLastError() == hipSuccess ? SILT_OK : SILT_EDEVICE;
}

// pond_ruts (src/fluid_core.rs:5509-5611); the pass books the FP64 change of
// the water it makes as a handoff line.
int silt_core_pond_ruts(silt_core* c, float* moved) {
  if (!c) return SILT_EINVAL;
  if (int rb = ensure_base(c)) return rb;
  if (restrict_fine(c)) return SILT_EDEVICE;
  core_pond_ruts(c->c, c->pond, c->d_pond, handoff_book(c), c->st);
  if (hipGetLastError() != hipSuccess) return SILT_EDEVICE;
  if (moved) {
    PondResult r{};
    HIPCHK(hipMemcpyAsync(&r, c->d_pond, sizeof r, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    *moved = r.moved;
  }
  return SILT_OK;
}

namespace {

int seat_launch(silt_core* c, int gate, float* eta, int* has_floor) {
  core_seat_rut_water(c->c, c->pond, gate, c->d_pond, handoff_book(c), c->st);
  if (hipGetLastError() != hipSuccess) return SILT_EDEVICE;
  if (eta || has_floor) {
    PondResult r{};
    HIPCHK(hipMemcpyAsync(&r, c->d_pond, sizeof r, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (eta) *eta = r.has_floor == 1 ? r.eta : NAN;
    if (has_floor) *has_floor = r.has_floor;
  }
  return SILT_OK;
}

}  // namespace

// seat_rut_water (src/fluid_core.rs:5622-5736) with the classify closure as
// host arrays.
int silt_core_seat_rut_water(silt_core* c, const float* pressed, const uint8_t* is_rut, float* eta, int* has_floor) {
  if (!c || !pressed || !is_rut) return SILT_EINVAL;
  if (int rb = ensure_base(c)) return rb;
  if (restrict_fine(c)) return SILT_EDEVICE;
  HIPCHK(hipMemcpyAsync(c->pond.f0, pressed, WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipMemcpyAsync(c->pond.m1, is_rut, WN, hipMemcpyHostToDevice, c->st));
  return seat_launch(c, 0, eta, has_floor);
}

// seat_rut_sheet (src/world.rs:5539-5561) on the core's ground: a window
// whose coarse sheet is dry (and whose patch is off) with no cut is skipped.
int silt_core_seat_rut_sheet(silt_core* c, int* seated) {
  if (!c) return SILT_EINVAL;
  if (!c->ground) return SILT_EINVAL;
  if (int rb = ensure_base(c)) return rb;
  const int gate = c->fine_on ? 0 : 1;
  if (restrict_fine(c)) return SILT_EDEVICE;
  core_seat_classify(c->soil_bed, c->soil_placed, c->soil_init, c->pond, c->st);
  int has = 0;
  int rc = seat_launch(c, gate, nullptr, seated ? &has : nullptr);
  if (rc) return rc;
  if (seated) *seated = has == 1 ? 1 : 0;
  return SILT_OK;
}

int silt_core_get_grain(silt_core* c, uint8_t* grain) {
  if (!c || !grain) return SILT_EINVAL;
  if (restrict_fine(c)) return SILT_EDEVICE;
  HIPCHK(hipMemcpyAsync(grain, c->c.grain, WN, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_get_suspended(silt_core* c, float* susp, double* accounts2) {
  if (!c) return SILT_EINVAL;
  if (susp) HIPCHK(hipMemcpyAsync(susp, c->sus.susp, WN * 4, hipMemcpyDeviceToHost, c->st));
  if (accounts2) HIPCHK(hipMemcpyAsync(accounts2, c->sus.acc, 2 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_set_suspended(silt_core* c, const float* susp) {
  if (!c || !susp) return SILT_EINVAL;
  HIPCHK(hipMemcpyAsync(c->sus.susp, susp, WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_get_tracers(silt_core* c, float* foam, float* turbid, double* accounts2) {
  if (!c) return SILT_EINVAL;
  if (foam) HIPCHK(hipMemcpyAsync(foam, c->tr.foam, WN * 4, hipMemcpyDeviceToHost, c->st));
  if (turbid) HIPCHK(hipMemcpyAsync(turbid, c->tr.turbid, WN * 4, hipMemcpyDeviceToHost, c->st));
  if (accounts2) HIPCHK(hipMemcpyAsync(accounts2, c->tr.acc, 2 * sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_set_tracers(silt_core* c, const float* foam, const float* turbid) {
  if (!c) return SILT_EINVAL;
  std::vector<float> f(WN);
  if (foam) {
    HIPCHK(hipMemcpyAsync(c->tr.foam, foam, WN * 4, hipMemcpyHostToDevice, c->st));
    memcpy(f.data(), foam, WN * 4);
  } else {
    HIPCHK(hipMemcpyAsync(f.data(), c->tr.foam, WN * 4, hipMemcpyDeviceToHost, c->st));
  }
  if (turbid) HIPCHK(hipMemcpyAsync(c->tr.turbid, turbid, WN * 4, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  std::vector<uint8_t> live(WN);
  for (int k = 0; k < WN; k++) live[k] = isfinite(f[k]) && f[k] > 1.0e-4f;
  HIPCHK(hipMemcpyAsync(c->tr.foam_live, live.data(), WN, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

int silt_core_ledger_full(silt_core* c, silt_fluid_ledger_full* out) {
  if (!out) return SILT_EINVAL;
  silt_fluid_ledger l{};
  int rc = silt_core_ledger(c, &l);
  if (rc) return rc;
  *out = silt_fluid_ledger_full{};
  out->water_m3 = l.water_m3;
  out->slurry_m3 = l.slurry_m3;
  out->frozen_water_m3 = l.frozen_water_m3;
  out->frozen_slurry_m3 = l.frozen_slurry_m3;
  out->spray_m3 = l.spray_m3;
  // The suspended-solid lines (:6830, :6840, :6876-6879): the column in the
  // water, what settled back, what is parked off the window, and the
  // entrained source; the total counts where the solids are.
  double acc[2];
  if ((rc = silt_core_get_suspended(c, nullptr, acc))) return rc;
  std::vector<float> sp(WN);
  HIPCHK(hipMemcpyAsync(sp.data(), c->sus.susp, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  double col = 0.0, parked = 0.0;
  for (float v : sp) col += (double)fmaxf(v, 0.0f);
  for (const auto& kv : c->parked) parked += (double)kv.second.susp;
  out->suspended_m3 = col * kCoarseArea;
  out->settled_suspended_m3 = acc[1];
  out->frozen_suspended_m3 = parked * kCoarseArea;
  out->entrained_m3 = acc[0];
  out->total_m3 = l.total_m3 + out->suspended_m3 + out->settled_suspended_m3 + out->frozen_suspended_m3;
  return SILT_OK;
}

}  // extern "C"

namespace {

// column_sum (src/fluid_core.rs:4617-4620) over a host plane: the
// reference's sequential f32 sum, so the volume restore_wiped puts back is
// the reference's to the bit.
float column_sum_host(const std::vector<float>& h) {
  float s = -0.0f;
  for (float x : h) s += fmaxf(x, 0.0f);
  return s * (kCoarseCell * kCoarseCell);
}

// restore_wiped (:4622-4637): volume a step lost to its dry cut goes back
// into the deepest cell (the last of equal ones, as max_by keeps).
void restore_wiped_host(std::vector<float>& h, float before) {
  const float lost = before - column_sum_host(h);
  if (lost <= 1.0e-8f) return;
  int best = -1;
  for (int k = 0; k < (int)h.size(); k++)
    if (isfinite(h[k]) && (best < 0 || !(h[k] < h[best]))) best = k;
  if (best >= 0) h[best] += lost / (kCoarseCell * kCoarseCell);
}

float peak_speed_host(const std::vector<float>& h, const std::vector<float>& u, const std::vector<float>& v) {
  float peak = 0.0f;
  for (size_t k = 0; k < h.size(); k++)
    if (h[k] > 1.0e-4f) peak = fmaxf(peak, fmaxf(fabsf(u[k]), fabsf(v[k])));
  return peak;
}

}  // namespace

extern "C" {

// drain_over_bed (src/fluid_core.rs:5123-5214): settle a pour with gravity
// steps (no confinement, pace 2, extra friction), restoring what the dry cut
// wipes, until the sheet has moved and then stayed calm for 24 steps.
float silt_core_drain_over_bed(silt_core* c, float dt, unsigned steps) {
  if (!c) return NAN;
  c->rebase = true;
  auto volume = [&]() -> float {
    std::vector<float> w(WN);
    if (hipMemcpyAsync(w.data(), c->c.water, WN * 4, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess)
      return NAN;
    return column_sum_host(w);
  };
  if (!(isfinite(dt) && dt > 0.0f) || steps == 0) return volume();
  if (restrict_fine(c)) return NAN;
  Workspace* ws = workspace_for(kWindowRes);
  if (!ws) return NAN;
  unsigned calm = 0;
  bool moved = false;
  const float keep = expf(-0.55f * dt);
  std::vector<float> hw(WN), hs(WN), a(WN), b(WN), d(WN), e(WN);
  for (unsigned it = 0; it < steps; it++) {
    if (hipMemcpyAsync(hw.data(), c->c.water, WN * 4, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
        hipMemcpyAsync(hs.data(), c->c.slurry, WN * 4, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess)
      return NAN;
    const float wb = column_sum_host(hw), sb = column_sum_host(hs);
    silt_sheet_opts o;
    silt_sheet_opts_default(&o);
    o.path = c->path;
    o.row_split = 1;
    o.pace = 2;
    o.sub_cap = 8.0f;
    o.banks = 1;
    SheetPlanes sp = ws->planes();
    core_soften(c->c.uw, c->c.vw, kWindowRes, c->st);
    sp.h = c->c.water;
    sp.u = c->c.uw;
    sp.v = c->c.vw;
    sp.bed = c->c.bed;
    silt_drive dg{SILT_DRIVE_GRAVITY, 0.0f, 0.0f, 0.0f};
    if (step_device_planes(sp, kWindowRes, kCoarseCell, dt, dg, o, true, true, dynamic_viscosity(3, 1.0f), c->st) < 0)
      return NAN;
    core_nu(c->nu_c, c->c.grain, c->c.moisture, WN, c->st);
    core_soften(c->c.us, c->c.vs, kWindowRes, c->st);
    sp.h = c->c.slurry;
    sp.u = c->c.us;
    sp.v = c->c.vs;
    sp.nu = c->nu_c;
    if (step_device_planes(sp, kWindowRes, kCoarseCell, dt, dg, o, true, false, 0.0f, c->st) < 0) return NAN;
    float* vel[4] = {c->c.uw, c->c.vw, c->c.us, c->c.vs};
    std::vector<float>* hv[4] = {&a, &b, &d, &e};
    for (int k = 0; k < 4; k++)
      if (hipMemcpyAsync(hv[k]->data(), vel[k], WN * 4, hipMemcpyDeviceToHost, c->st) != hipSuccess) return NAN;
    if (hipMemcpyAsync(hw.data(), c->c.water, WN * 4, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
        hipMemcpyAsync(hs.data(), c->c.slurry, WN * 4, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess)
      return NAN;
    for (int k = 0; k < 4; k++)
      for (float& x : *hv[k]) x *= keep;
    restore_wiped_host(hw, wb);
    restore_wiped_host(hs, sb);
    for (int k = 0; k < 4; k++)
      if (hipMemcpyAsync(vel[k], hv[k]->data(), WN * 4, hipMemcpyHostToDevice, c->st) != hipSuccess) return NAN;
    if (hipMemcpyAsync(c->c.water, hw.data(), WN * 4, hipMemcpyHostToDevice, c->st) != hipSuccess ||
        hipMemcpyAsync(c->c.slurry, hs.data(), WN * 4, hipMemcpyHostToDevice, c->st) != hipSuccess ||
        hipStreamSynchronize(c->st) != hipSuccess)
      return NAN;
    const float speed = fmaxf(peak_speed_host(hw, a, b), peak_speed_host(hs, d, e));
    if (speed > 0.05f) moved = true;
    if (moved && speed < 0.008f) {
      if (++calm >= 24) break;
    } else {
      calm = 0;
    }
  }
  return volume();
}
int silt_surrogate_targets_debug(const float* hn_w, const float* hc_w, const float* hn_s, const float* hc_s, int i0,
                                 int j0, int nx, int nz, float* targets) {
  if (!hn_w || !hc_w || !hn_s || !hc_s || !targets || nx < 2 || nz < 2 || nx > kSurMax || nz > kSurMax || i0 < 0 ||
      j0 < 0 || i0 + nx > kWindowRes || j0 + nz > kWindowRes)
    return SILT_EINVAL;
  float* d = nullptr;
  if (hipMalloc(&d, (4 * (size_t)WN + kSurTgtFloats) * sizeof(float)) != hipSuccess) return SILT_ENOMEM;
  const float* src[4] = {hn_w, hc_w, hn_s, hc_s};
  int rc = SILT_OK;
  for (int k = 0; k < 4 && rc == SILT_OK; k++)
    if (hipMemcpy(d + k * WN, src[k], WN * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) rc = SILT_EDEVICE;
  if (rc == SILT_OK) {
    sur_targets(d, d + WN, d + 2 * WN, d + 3 * WN, kWindowRes, SurRegion{i0, j0, nx, nz}, d + 4 * WN, nullptr);
    if (hipMemcpy(targets, d + 4 * WN, (size_t)nx * nz * 8 * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      rc = SILT_EDEVICE;
  }
  (void)hipFree(d);
  return rc;
}

int silt_surrogate_apply_debug(float* water, float* slurry, const float* y, int i0, int j0, int nx, int nz) {
  if (!water || !slurry || !y || nx < 2 || nz < 2 || nx > kSurMax || nz > kSurMax || i0 < 0 || j0 < 0 ||
      i0 + nx > kWindowRes || j0 + nz > kWindowRes)
    return SILT_EINVAL;
  float* d = nullptr;
  const size_t ny = (size_t)nx * nz * 8;
  if (hipMalloc(&d, (2 * (size_t)WN + ny + 2 * kSurMax * kSurMax) * sizeof(float)) != hipSuccess) return SILT_ENOMEM;
  int rc = SILT_OK;
  if (hipMemcpy(d, water, WN * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + WN, slurry, WN * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + 2 * WN, y, ny * 4, hipMemcpyHostToDevice) != hipSuccess)
    rc = SILT_EDEVICE;
  if (rc == SILT_OK) {
    sur_apply(d, d + WN, kWindowRes, d + 2 * WN, SurRegion{i0, j0, nx, nz}, d + 2 * WN + ny, nullptr);
    if (hipMemcpy(water, d, WN * 4, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(slurry, d + WN, WN * 4, hipMemcpyDeviceToHost) != hipSuccess)
      rc = SILT_EDEVICE;
  }
  (void)hipFree(d);
  return rc;
}

}  // extern "C"

namespace {

// world_key (src/fluid_core.rs:9894-9896) packed into 64 bits.
uint64_t park_key(float x, float z) {
  const int kx = (int)floorf(x / kCoarseCell), kz = (int)floorf(z / kCoarseCell);
  return ((uint64_t)(uint32_t)kx << 32) | (uint64_t)(uint32_t)kz;
}

// park_slot (src/fluid_core.rs:4797-4805) for the planes this library holds.
// park_slot (src/fluid_core.rs:4797-4805) without the dust and suspended
// columns this library does not carry.
bool park_worth(const silt_core::Parked& p, float ambient) {
  return p.water > 1.0e-8f || p.slurry > 1.0e-8f || p.susp > 1.0e-8f || fabsf(p.moisture - ambient) > 6.0e-3f ||
         p.foam > 1.0e-4f || p.turbid > 1.0e-8f;
}

// each_leaving (src/fluid_core.rs:9936-9986): cells that leave the window
// when its origin moves by (di, dj) cells, in the reference's order.
template <typename Fn>
void each_leaving(int di, int dj, Fn f) {
  const int n = kWindowRes;
  if (di > 0) {
    const int w = std::min(di, n);
    for (int j = 0; j < n; j++)
      for (int i = 0; i < w; i++) f(i, j);
  } else if (di < 0) {
    const int w = std::min(-di, n);
    for (int j = 0; j < n; j++)
      for (int i = n - w; i < n; i++) f(i, j);
  }
  auto covered = [&](int i) { return di > 0 ? i < di : di < 0 ? i >= n + di : false; };
  if (dj > 0) {
    const int h = std::min(dj, n);
    for (int j = 0; j < h; j++)
      for (int i = 0; i < n; i++)
        if (!covered(i)) f(i, j);
  } else if (dj < 0) {
    const int h = std::min(-dj, n);
    for (int j = n - h; j < n; j++)
      for (int i = 0; i < n; i++)
        if (!covered(i)) f(i, j);
  }
}

// uniform_cell_shift (src/fluid_core.rs:9899-9934).
bool uniform_cell_shift(const float* o, const float* nw, int* di, int* dj) {
  auto key = [](const float* org, int i, int j, int* kx, int* kz) {
    *kx = (int)floorf((org[0] + ((float)i + 0.5f) * kCoarseCell) / kCoarseCell);
    *kz = (int)floorf((org[1] + ((float)j + 0.5f) * kCoarseCell) / kCoarseCell);
  };
  int o0x, o0z, n0x, n0z, o1x, o1z, n1x, n1z, omx, omz, nmx, nmz;
  const int last = kWindowRes - 1, mid = kWindowRes / 2;
  key(o, 0, 0, &o0x, &o0z);
  key(nw, 0, 0, &n0x, &n0z);
  *di = n0x - o0x;
  *dj = n0z - o0z;
  key(o, last, last, &o1x, &o1z);
  key(nw, last, last, &n1x, &n1z);
  if (o1x - o0x != last || o1z - o0z != last) return false;
  if (n1x - n0x != last || n1z - n0z != last) return false;
  if (n1x - o1x != *di || n1z - o1z != *dj) return false;
  key(o, mid, mid, &omx, &omz);
  key(nw, mid, mid, &nmx, &nmz);
  return nmx - omx == *di && nmz - omz == *dj;
}

}  // namespace

extern "C" {

// recenter (src/fluid_core.rs:5409-5422) with shift_coarse and its scroll
// and hash forms (:9004-9226): cells that slide off are parked by world cell,
// cells that slide on come back from the parked set; bed and grain stay on
// their array index (the caller reads the substrate again).
int silt_core_recenter(silt_core* c, float x, float z) {
  if (!c || !isfinite(x) || !isfinite(z)) return SILT_EINVAL;
  c->rebase = true;
  if (restrict_fine(c)) return SILT_EDEVICE;
  c->center[0] = x;
  c->center[1] = z;
  const float extent = (float)kWindowRes * kCoarseCell;
  const float nw[2] = {floorf((x - extent * 0.5f) / kCoarseCell) * kCoarseCell,
                       floorf((z - extent * 0.5f) / kCoarseCell) * kCoarseCell};
  if (fabsf(nw[0] - c->origin[0]) < 1.0e-4f && fabsf(nw[1] - c->origin[1]) < 1.0e-4f) return SILT_OK;
  float* dev[10] = {c->c.water, c->c.slurry, c->c.uw, c->c.vw, c->c.us, c->c.vs, c->c.moisture, c->tr.foam,
                    c->tr.turbid, c->sus.susp};
  std::vector<float> h[10], o[10];
  for (int k = 0; k < 10; k++) {
    h[k].resize(WN);
    HIPCHK(hipMemcpyAsync(h[k].data(), dev[k], WN * 4, hipMemcpyDeviceToHost, c->st));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  const float amb = c->moisture_ambient;
  auto slot_at = [&](int idx) {
    return silt_core::Parked{h[0][idx], h[1][idx], h[2][idx], h[3][idx], h[4][idx],
                             h[5][idx], h[6][idx], h[7][idx], h[8][idx], h[9][idx]};
  };
  auto park = [&](const float* org, int i, int j) {
    const int idx = j * kWindowRes + i;
    const uint64_t key = park_key(org[0] + ((float)i + 0.5f) * kCoarseCell, org[1] + ((float)j + 0.5f) * kCoarseCell);
    const silt_core::Parked p = slot_at(idx);
    if (park_worth(p, amb))
      c->parked[key] = p;
    else
      c->parked.erase(key);
  };
  auto restore = [&](const float* org, int i, int j) {
    const int idx = j * kWindowRes + i;
    const uint64_t key = park_key(org[0] + ((float)i + 0.5f) * kCoarseCell, org[1] + ((float)j + 0.5f) * kCoarseCell);
    auto it = c->parked.find(key);
    if (it == c->parked.end()) return;
    const silt_core::Parked& p = it->second;
    const float v[10] = {p.water, p.slurry, p.uw, p.vw, p.us, p.vs, p.moisture, p.foam, p.turbid, p.susp};
    for (int k = 0; k < 10; k++) o[k][idx] = v[k];
    c->parked.erase(it);
  };
  for (int k = 0; k < 10; k++) o[k].assign(WN, k == 6 ? amb : 0.0f);
  int di = 0, dj = 0;
  if (uniform_cell_shift(c->origin, nw, &di, &dj)) {
    const float old[2] = {c->origin[0], c->origin[1]};
    each_leaving(di, dj, [&](int i, int j) { park(old, i, j); });
    for (int j = 0; j < kWindowRes; j++)
      for (int i = 0; i < kWindowRes; i++) {
        const int si = i + di, sj = j + dj;
        if (si < 0 || sj < 0 || si >= kWindowRes || sj >= kWindowRes) continue;
        const int src = sj * kWindowRes + si, dst = j * kWindowRes + i;
        if (!(h[0][src] > 1.0e-8f || h[1][src] > 1.0e-8f || h[6][src] > 1.0e-4f || h[7][src] > 1.0e-5f ||
              h[8][src] > 1.0e-8f || h[9][src] > 1.0e-8f))
          continue;
        for (int k = 0; k < 10; k++) o[k][dst] = h[k][src];
      }
    c->origin[0] = nw[0];
    c->origin[1] = nw[1];
    each_leaving(-di, -dj, [&](int i, int j) { restore(c->origin, i, j); });
  } else {
    const float old[2] = {c->origin[0], c->origin[1]};
    for (int j = 0; j < kWindowRes; j++)
      for (int i = 0; i < kWindowRes; i++) park(old, i, j);
    c->origin[0] = nw[0];
    c->origin[1] = nw[1];
    for (int j = 0; j < kWindowRes; j++)
      for (int i = 0; i < kWindowRes; i++) restore(c->origin, i, j);
  }
  for (int k = 0; k < 10; k++) HIPCHK(hipMemcpyAsync(dev[k], o[k].data(), WN * 4, hipMemcpyHostToDevice, c->st));
  // reset_tracer_indices (:9228-9233): no foam is live after a move.
  core_fill_bytes(c->tr.foam_live, 0, WN, c->st);
  HIPCHK(hipStreamSynchronize(c->st));
  return SILT_OK;
}

// lay_slurry_on_water (src/fluid_core.rs:5236-5254): slurry of `depth` on
// every cell holding more than 12 mm of water; returns the volume laid (the
// reference's f32 sum).
float silt_core_lay_slurry_on_water(silt_core* c, float depth, int grain) {
  if (!c || !(isfinite(depth) && depth > 0.0f) || grain < 0 || grain > 4) return 0.0f;
  c->rebase = true;
  if (restrict_fine(c)) return NAN;
  std::vector<float> w(WN), sl(WN), m(WN);
  std::vector<uint8_t> g(WN);
  if (hipMemcpyAsync(w.data(), c->c.water, WN * 4, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipMemcpyAsync(sl.data(), c->c.slurry, WN * 4, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipMemcpyAsync(m.data(), c->c.moisture, WN * 4, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipMemcpyAsync(g.data(), c->c.grain, WN, hipMemcpyDeviceToHost, c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return NAN;
  const float area = kCoarseCell * kCoarseCell;
  float vol = 0.0f;
  for (int k = 0; k < WN; k++) {
    if (w[k] <= 0.012f) continue;
    sl[k] = depth;
    g[k] = (uint8_t)grain;
    m[k] = fmaxf(m[k], 0.90f);
    vol += depth * area;
  }
  if (hipMemcpyAsync(c->c.slurry, sl.data(), WN * 4, hipMemcpyHostToDevice, c->st) != hipSuccess ||
      hipMemcpyAsync(c->c.moisture, m.data(), WN * 4, hipMemcpyHostToDevice, c->st) != hipSuccess ||
      hipMemcpyAsync(c->c.grain, g.data(), WN, hipMemcpyHostToDevice, c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess)
    return NAN;
  return vol;
}

// window_bounds (src/fluid_core.rs:6977-6980): corner x, z and the extent.
int silt_core_window_bounds(const silt_core* c, float* out3) {
  if (!c || !out3) return SILT_EINVAL;
  out3[0] = c->origin[0];
  out3[1] = c->origin[1];
  out3[2] = (float)kWindowRes * kCoarseCell;
  return SILT_OK;
}

long silt_core_parked_cells(const silt_core* c) { return c ? (long)c->parked.size() : SILT_EINVAL; }

}  // extern "C"

// Training set and trainer of the tire sub-grid network (include/silt_mlp.h).
struct silt_trainer {
  size_t max_steps = 0;
  unsigned seed = 0;
  silt_core* twin = nullptr;
  float* ctx = nullptr;  // [max_steps][kSurCtxFloats]
  float* tgt = nullptr;  // [max_steps][kSurTgtFloats]
  SurRegion* regions = nullptr;
  std::vector<SurRegion> h_regions;
  size_t nsteps = 0;
  Mlp net;
  hipStream_t st = nullptr;
  uint32_t iter = 0;
};

extern "C" {

silt_trainer* silt_trainer_create(size_t max_steps, unsigned seed) {
  if (max_steps == 0 || max_steps > (1u << 20)) return nullptr;
  silt_trainer* t = new silt_trainer();
  t->max_steps = max_steps;
  t->seed = seed;
  bool ok = hipStreamCreateWithFlags(&t->st, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc(&t->ctx, max_steps * kSurCtxFloats * sizeof(float)) == hipSuccess &&
            hipMalloc(&t->tgt, max_steps * kSurTgtFloats * sizeof(float)) == hipSuccess &&
            hipMalloc(&t->regions, max_steps * sizeof(SurRegion)) == hipSuccess &&
            mlp_init(t->net, 512, seed, t->st) == 0 && hipStreamSynchronize(t->st) == hipSuccess;
  if (!ok) {
    silt_trainer_destroy(t);
    return nullptr;
  }
  t->h_regions.resize(max_steps);
  return t;
}

void silt_trainer_destroy(silt_trainer* t) {
  if (!t) return;
  if (t->st) (void)hipStreamSynchronize(t->st);
  mlp_free(t->net);
  if (t->twin) silt_core_destroy(t->twin);
  if (t->ctx) (void)hipFree(t->ctx);
  if (t->tgt) (void)hipFree(t->tgt);
  if (t->regions) (void)hipFree(t->regions);
  if (t->st) (void)hipStreamDestroy(t->st);
  delete t;
}

int silt_trainer_record(silt_trainer* t, silt_core* c, float dt, silt_spray_event* events, size_t cap,
                        size_t* nevents) {
  if (!t || !c) return SILT_EINVAL;
  SurRegion r = sur_region(c, c->pushes);
  if (!(dt > 0.0f) || !r.nx || t->nsteps >= t->max_steps || c->tire_model != SILT_TIRE_NESTED) {
    int rc = silt_core_step(c, dt, events, cap, nevents);
    return rc < 0 ? rc : 0;
  }
  if (!t->twin) {
    t->twin = silt_core_create(c->origin[0] + 0.5f * kWindowRes * kCoarseCell,
                               c->origin[1] + 0.5f * kWindowRes * kCoarseCell, c->path);
    if (!t->twin) return SILT_ENOMEM;
  }
  silt_core* w = t->twin;
  int rc = restrict_fine(c);
  if (rc) return rc;
  // The twin takes the core's state and tires and steps coarse-only.
  w->origin[0] = c->origin[0];
  w->origin[1] = c->origin[1];
  w->coarse_confine = c->coarse_confine;
  w->near_on = false;
  w->tire_model = SILT_TIRE_COARSE;
  float* src[8] = {c->c.water, c->c.slurry, c->c.uw, c->c.vw, c->c.us, c->c.vs, c->c.bed, c->c.moisture};
  float* dst[8] = {w->c.water, w->c.slurry, w->c.uw, w->c.vw, w->c.us, w->c.vs, w->c.bed, w->c.moisture};
  for (int k = 0; k < 8; k++)
    HIPCHK(hipMemcpyAsync(dst[k], src[k], WN * sizeof(float), hipMemcpyDeviceToDevice, c->st));
  HIPCHK(hipMemcpyAsync(w->c.grain, c->c.grain, WN, hipMemcpyDeviceToDevice, c->st));
  w->rebase = true;
  const PushList pl = push_list(c->pushes);
  sur_context(core_grid(c), r, pl, t->ctx + t->nsteps * kSurCtxFloats, c->st);
  HIPCHK(hipStreamSynchronize(c->st));
  w->pushes = c->pushes;
  if ((rc = silt_core_step(w, dt, nullptr, 0, nullptr)) < 0) return rc;
  if ((rc = silt_core_step(c, dt, events, cap, nevents)) < 0) return rc;
  sur_targets(c->c.water, w->c.water, c->c.slurry, w->c.slurry, kWindowRes, r, t->tgt + t->nsteps * kSurTgtFloats,
              c->st);
  t->h_regions[t->nsteps] = r;
  HIPCHK(hipMemcpyAsync(t->regions + t->nsteps, &t->h_regions[t->nsteps], sizeof(SurRegion),
                        hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  t->nsteps++;
  return (int)t->nsteps;
}

long silt_trainer_steps(const silt_trainer* t) { return t ? (long)t->nsteps : SILT_EINVAL; }

long silt_trainer_samples(const silt_trainer* t) {
  if (!t) return SILT_EINVAL;
  long n = 0;
  for (size_t s = 0; s < t->nsteps; s++) n += (long)t->h_regions[s].nx * t->h_regions[s].nz;
  return n;
}

int silt_trainer_fit(silt_trainer* t, size_t train_steps, int iterations, int batch, float lr, double* last_mse) {
  if (!t || iterations < 0 || batch < 1 || batch > 8192 || !(lr > 0.0f)) return SILT_EINVAL;
  if (train_steps == 0 || train_steps > t->nsteps) train_steps = t->nsteps;
  if (train_steps == 0) return SILT_EINVAL;
  int rc = mlp_reserve(t->net, (size_t)batch);
  if (rc) return rc;
  for (int it = 0; it < iterations; it++) {
    sur_gather_random(t->ctx, t->tgt, t->regions, (int)train_steps, t->seed * 2654435761u + ++t->iter, batch,
                      t->net.xb, t->net.tg, t->st);
    mlp_forward(t->net, (size_t)batch);
    mlp_backward(t->net, (size_t)batch);
    mlp_adam(t->net, lr);
  }
  double sse = 0.0;
  HIPCHK(hipMemcpyAsync(&sse, t->net.loss, sizeof sse, hipMemcpyDeviceToHost, t->st));
  HIPCHK(hipStreamSynchronize(t->st));
  if (last_mse) *last_mse = iterations ? sse / ((double)batch * kMlpOut) : 0.0;
  return SILT_OK;
}

int silt_trainer_eval(silt_trainer* t, size_t first, size_t count, double* mse_net, double* mse_zero) {
  if (!t || first + count > t->nsteps || count == 0) return SILT_EINVAL;
  int rc = mlp_reserve(t->net, (size_t)kSurMax * kSurMax);
  if (rc) return rc;
  double se = 0.0, sz = 0.0;
  size_t n = 0;
  std::vector<float> y((size_t)kSurMax * kSurMax * kMlpOut), tg(y.size());
  for (size_t s = first; s < first + count; s++) {
    const SurRegion& r = t->h_regions[s];
    const size_t cells = (size_t)r.nx * r.nz;
    sur_gather_region(t->ctx + s * kSurCtxFloats, r, t->net.xb, t->st);
    mlp_infer(t->net, cells);
    HIPCHK(hipMemcpyAsync(y.data(), t->net.y, cells * kMlpOut * sizeof(float), hipMemcpyDeviceToHost, t->st));
    HIPCHK(hipMemcpyAsync(tg.data(), t->tgt + s * kSurTgtFloats, cells * kMlpOut * sizeof(float),
                          hipMemcpyDeviceToHost, t->st));
    HIPCHK(hipStreamSynchronize(t->st));
    for (size_t k = 0; k < cells * kMlpOut; k++) {
      const double d = (double)y[k] - tg[k];
      se += d * d;
      sz += (double)tg[k] * tg[k];
    }
    n += cells * kMlpOut;
  }
  if (mse_net) *mse_net = se / (double)n;
  if (mse_zero) *mse_zero = sz / (double)n;
  return SILT_OK;
}

int silt_trainer_params(const silt_trainer* t, float* params) {
  if (!t || !params) return SILT_EINVAL;
  HIPCHK(hipMemcpyAsync(params, t->net.master, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, t->st));
  HIPCHK(hipStreamSynchronize(t->st));
  return SILT_OK;
}

int silt_trainer_set_params(silt_trainer* t, const float* params) {
  if (!t || !params) return SILT_EINVAL;
  HIPCHK(hipMemcpyAsync(t->net.master, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, t->st));
  mlp_sync_bf16(t->net);
  HIPCHK(hipStreamSynchronize(t->st));
  return SILT_OK;
}

}  // extern "C"
