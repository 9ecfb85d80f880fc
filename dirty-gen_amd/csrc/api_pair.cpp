// Two-layer sheet handle (include/silt_pair.h): water and slurry over one
// bed, each stepped as the reference's generic-res step_sheet
// (src/fluid_core.rs:770-785) once per step, with the layer constants of
// FluidCore::step_pair (:8036-8125): water Gravity{0.08} at the water
// viscosity, slurry Gravity{0.05} at its own viscosity.  State stays on the
// GPU between steps; the plan of each step is read from the statistics the
// previous step's kernel left behind (a 64-byte read back), and the ledger is
// kept in FP64 with explicit dry-cut and residue lines.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "core.hpp"
#include "mlp.hpp"
#include "plan.hpp"
#include "rut.hpp"
#include "sheet_device.hpp"
#include "sheet_fused.hpp"
#include "silt_mlp.h"
#include "silt_pair.h"
#include "silt_sheet.h"
#include "surrogate.hpp"
#include "workspace.hpp"

#define HIPCHK(x)                                 \
  do {                                            \
    if ((x) != hipSuccess) return SILT_EDEVICE;   \
  } while (0)

using namespace silt;

struct silt_pair {
  int res = 0;
  float cell = 0.2f;
  int path = SILT_PATH_FAST;
  int device = 0;
  size_t n = 0;
  float* bed = nullptr;
  float* planes[2][2][3] = {};  // [layer][buffer][h, u, v]
  int cur[2] = {0, 0};
  float* nu_slurry = nullptr;
  bool nu_slurry_uniform = true;
  float nu_slurry0 = 0.5f;
  float nu_water = 0.9f;  // dynamic_viscosity(Grain::Water, 1.0), src/fluid_core.rs:642-646
  float vort[2] = {0.08f, 0.05f};
  SheetStats* stats = nullptr;     // [2] device
  SheetTotals* totals = nullptr;   // [2] device
  SheetStats* h_stats = nullptr;   // pinned host copy
  double vol0[2] = {0, 0};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_kernel_ms = 0.0f;
  bool kernel_ms_pending = false;  // ev0 / ev1 recorded, time not read yet
  // Speculative launch of the last step's plan (silt_pair_step).
  bool speculate = true;
  bool spec_valid = false;
  float spec_dt = 0.0f;
  StepPlan spec_plans[2];
  SheetTotals* totals_save = nullptr;
  SheetStats* stats_save = nullptr;  // [2] device: stats before a speculative launch
  hipEvent_t ev_stats = nullptr;
  // The statistics' plan fields (wet count and box, largest speed) no longer
  // describe the planes (the network's fluxes moved depth after the step
  // wrote them): recomputed before a plan reads them.  The volume stays
  // exact (the flux apply adds its FP64 change), so the ledger needs none.
  bool stats_stale = false;
  long long spec_hits = 0, spec_misses = 0;
  // Back-off after an undone speculation: the next spec_pause steps plan
  // first and launch after; each miss doubles the pause (1, 2, 4 ... 64
  // steps), a speculation that holds resets it.
  int spec_pause = 0, spec_pause_next = 1;
  // XCD balance of the fused launches (silt_pair_balance): rows of each
  // tile scaled by its XCD's weight when xcd_on; per-workgroup clocks
  // recorded into d_clocks while calibrating.
  float xcd_w[kXcds] = {1, 1, 1, 1, 1, 1, 1, 1};
  bool xcd_on = false;
  unsigned long long* d_clocks = nullptr;
  bool clocks_on = false;
  std::vector<unsigned long long> h_clocks;  // the last calibration step's records
  int last_nsub[2] = {0, 0};
  int last_fallback = 0;
  PlanLabeller lab;
  PlanScratch ps[2];
  // Tires on the sheet (silt_pair_step_tires): the ground the ruts cut (the
  // sheet bed is ground bed + placed), the world corner of cell (0, 0), the
  // network and its per-vehicle context and limiter scratch.
  float origin[2] = {0, 0};
  float *soil_bed = nullptr, *soil_placed = nullptr, *soil_init = nullptr, *soil_rut = nullptr;
  RutContact* d_contacts = nullptr;
  size_t cap_contacts = 0;
  bool net_ready = false;
  Mlp net;
  float* sur_ctx = nullptr;
  float* sur_fac = nullptr;
  size_t cap_groups = 0;
  // Per-vehicle launch table (TireJob records, then the pushes) and its
  // pinned staging copy.
  void* d_tires = nullptr;
  void* h_tires = nullptr;
  size_t cap_tires = 0;
  bool tire_batching = true;
  hipEvent_t tev[4] = {nullptr, nullptr, nullptr, nullptr};
  float last_tire_ms = 0.0f;
};

namespace {

const float kEps = 1.0e-6f;

SheetPlanes layer_planes(silt_pair* p, int layer, int buf) {
  SheetPlanes s{};
  s.h = p->planes[layer][buf][0];
  s.u = p->planes[layer][buf][1];
  s.v = p->planes[layer][buf][2];
  s.bed = p->bed;
  s.nu = layer == 1 && !p->nu_slurry_uniform ? p->nu_slurry : nullptr;
  s.gate = nullptr;
  return s;
}

float layer_nu0(const silt_pair* p, int layer) { return layer == 0 ? p->nu_water : p->nu_slurry0; }
bool layer_nu_uniform(const silt_pair* p, int layer) { return layer == 0 || p->nu_slurry_uniform; }

// The plan of both layers from the statistics in p->h_stats (of the planes
// p->cur points at): boxes from one batched labelling, substep counts with
// at most one more wait.
int pair_plan(silt_pair* p, float dt, StepPlan* plans) {
  hipStream_t st = p->stream;
  const int res = p->res;
  p->last_fallback = 0;
  // Both layers' boxes in one labelling, both substep counts with one wait.
  PlanSheet sheets[2];
  char host[2];
  for (int l = 0; l < 2; l++) sheets[l] = PlanSheet{p->planes[l][p->cur[l]][0], res, p->cell, p->h_stats + l};
  int pr = plan_boxes_many(p->lab, sheets, 2, dt, kEps, plans, host, st);
  if (pr < 0) return pr;
  bool wait = false;
  for (int l = 0; l < 2; l++) {
    float** pl = p->planes[l][p->cur[l]];
    if (host[l]) {
      // A 2 cm sheet (sleeping tiles) or more wet clusters than are read
      // back: the planes go to the host.
      p->last_fallback |= 1 << l;
      std::vector<float> hh(p->n), uh(p->n), vh(p->n), bh(p->n);
      size_t fb = p->n * sizeof(float);
      HIPCHK(hipMemcpyAsync(hh.data(), pl[0], fb, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(uh.data(), pl[1], fb, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(vh.data(), pl[2], fb, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(bh.data(), p->bed, fb, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      plans[l] = plan_step(hh.data(), uh.data(), vh.data(), bh.data(), res, p->cell, dt, 0, 0.0f, 0.0f, kEps, 8.0f, 0);
      continue;
    }
    int rc = plan_nsub_begin(pl[1], pl[2], res, p->cell, dt, p->h_stats[l], p->ps[l], &plans[l], st);
    if (rc < 0) return rc;
    wait = wait || p->ps[l].maxs2_pending;
  }
  if (wait) HIPCHK(hipStreamSynchronize(st));
  for (int l = 0; l < 2; l++)
    if (!host[l]) plan_nsub_end(p->ps[l], p->cell, dt, &plans[l]);
  return 0;
}

// Enqueue the step of both layers for the given plans: fused launches
// (ping-pong into the other planes) and staged layers (in place).  Returns
// the largest substep count or a SILT_E* code.
int pair_launch(silt_pair* p, float dt, const StepPlan* plans) {
  hipStream_t st = p->stream;
  const int res = p->res;
  silt_sheet_opts o;
  silt_sheet_opts_default(&o);
  o.path = p->path;
  HIPCHK(hipEventRecord(p->ev0, st));
  // Layers whose plan the fused kernel covers and that take one substep run
  // together in one launch; anything else runs per layer.
  auto fused_ok = [&](int l) {
    if (p->path == SILT_PATH_EXACT || plans[l].boxes.empty() || plans[l].pace3) return false;
    if (!fused_boxes_fit(plans[l].boxes)) return false;
    for (const Box& b : plans[l].boxes)
      if (!reference_uses_jacobi(b, p->cell, 0, true)) return false;
    return true;
  };
  // The layers are independent: substep k of every fused layer that has one
  // shares a launch (a layer of one substep rides with the first substep of
  // the other).
  int fused[2], nfused = 0, most = 0;
  bool staged = false;
  for (int l = 0; l < 2; l++) {
    p->last_nsub[l] = plans[l].boxes.empty() ? 0 : plans[l].nsub;
    if (plans[l].boxes.empty()) continue;
    if (fused_ok(l)) {
      fused[nfused++] = l;
      most = std::max(most, plans[l].nsub);
      continue;
    }
    // Stage kernels in place on the current planes (Gauss-Seidel boxes,
    // overlapping boxes, or the exact path).
    silt_drive d{SILT_DRIVE_GRAVITY, 0.0f, 0.0f, p->vort[l]};
    Workspace* ws = workspace_for(res);
    if (!ws) return SILT_ENOMEM;
    SheetPlanes s = ws->planes();
    SheetPlanes cur = layer_planes(p, l, p->cur[l]);
    s.h = cur.h;
    s.u = cur.u;
    s.v = cur.v;
    s.bed = p->bed;
    s.nu = cur.nu ? cur.nu : s.nu;
    o.path = SILT_PATH_EXACT;
    int rc = run_plan(s, plans[l], res, p->cell, dt, d, o, true, layer_nu_uniform(p, l), layer_nu0(p, l), st,
                      p->totals + l);
    if (rc < 0) return rc;
    o.path = p->path;
    stats_compute(cur.h, cur.u, cur.v, res, kEps, p->stats + l, st);
    staged = true;
  }
  // Time the fused launches alone when they are the step's only kernel work:
  // one launch carries the events in its own dispatch (they time the kernel
  // itself), several are bracketed by event records.
  const bool one_launch = !staged && nfused && most == 1;
  if (!staged && nfused && !one_launch) HIPCHK(hipEventRecord(p->ev0, st));
  for (int k = 0; k < most; k++) {
    FusedLayer launch[2];
    int flip[2], m = 0;
    for (int q = 0; q < nfused; q++) {
      const int l = fused[q];
      if (plans[l].nsub <= k) continue;
      silt_drive d{SILT_DRIVE_GRAVITY, 0.0f, 0.0f, p->vort[l]};
      FusedLayer& f = launch[m++];
      f = FusedLayer{};
      f.in = layer_planes(p, l, p->cur[l]);
      const int nb = 1 - p->cur[l];
      f.ho = p->planes[l][nb][0];
      f.uo = p->planes[l][nb][1];
      f.vo = p->planes[l][nb][2];
      f.a = make_args(res, p->cell, dt / (float)plans[l].nsub, d, kEps, 0, true, layer_nu_uniform(p, l),
                      layer_nu0(p, l));
      f.boxes = plans[l].boxes;
      f.stats = p->stats + l;
      f.totals = p->totals + l;
      stats_reset(f.stats, 1, st);
      flip[m - 1] = l;
    }
    FusedBalance bal;
    bal.xcd_w = p->xcd_on ? p->xcd_w : nullptr;
    bal.clocks = p->clocks_on ? p->d_clocks : nullptr;
    if (one_launch) {
      bal.ev_start = p->ev0;
      bal.ev_stop = p->ev1;
    }
    int rc = fused_substep_layers(launch, m, p->bed, res, p->path == SILT_PATH_FUSED_EXACT, st, &bal);
    if (rc != 0) return rc;
    // The output planes become current only once their launch is enqueued.
    for (int q = 0; q < m; q++) p->cur[flip[q]] = 1 - p->cur[flip[q]];
  }
  if (!one_launch) HIPCHK(hipEventRecord(p->ev1, st));
  p->kernel_ms_pending = true;
  // Staged layers use the process-wide workspace of this side length, which
  // the next call of another handle may use on its own stream: finish them
  // here.  Fused launches touch only this handle's buffers and may run on.
  if (staged) HIPCHK(hipStreamSynchronize(st));
  return std::max(p->last_nsub[0], p->last_nsub[1]);
}

// Whether a plan may be launched before it is confirmed: every layer fused
// with one substep, so the step only writes the other planes and the
// current ones stay as they were.
bool pair_speculable(const silt_pair* p, const StepPlan* plans) {
  if (p->path == SILT_PATH_EXACT) return false;
  for (int l = 0; l < 2; l++) {
    if (plans[l].boxes.empty()) continue;
    if (plans[l].nsub != 1 || plans[l].pace3 || !fused_boxes_fit(plans[l].boxes)) return false;
    for (const Box& b : plans[l].boxes)
      if (!reference_uses_jacobi(b, p->cell, 0, true)) return false;
  }
  return true;
}

bool same_plans(const StepPlan* a, const StepPlan* b) {
  for (int l = 0; l < 2; l++) {
    if (a[l].nsub != b[l].nsub || a[l].pace3 != b[l].pace3 || a[l].boxes.size() != b[l].boxes.size()) return false;
    for (size_t k = 0; k < a[l].boxes.size(); k++) {
      const Box &x = a[l].boxes[k], &y = b[l].boxes[k];
      if (x.i0 != y.i0 || x.i1 != y.i1 || x.j0 != y.j0 || x.j1 != y.j1) return false;
    }
  }
  return true;
}

}  // namespace

extern "C" {

silt_pair* silt_pair_create(size_t res, float cell, int path) {
  if (res < 1 || res > 65536 || !(cell > 0.0f)) return nullptr;
  if (path != SILT_PATH_EXACT && path != SILT_PATH_FAST && path != SILT_PATH_FUSED_EXACT) return nullptr;
  silt_pair* p = new silt_pair();
  p->res = (int)res;
  p->cell = cell;
  p->path = path;
  p->n = res * res;
  size_t fb = p->n * sizeof(float);
  bool ok = hipGetDevice(&p->device) == hipSuccess && hipMalloc(&p->bed, fb) == hipSuccess &&
            hipMalloc(&p->nu_slurry, fb) == hipSuccess;
  for (int l = 0; l < 2 && ok; l++)
    for (int b = 0; b < 2 && ok; b++)
      for (int k = 0; k < 3 && ok; k++) ok = hipMalloc(&p->planes[l][b][k], fb) == hipSuccess;
  ok = ok && hipMalloc(&p->stats, 2 * sizeof(SheetStats)) == hipSuccess &&
       hipMalloc(&p->totals, 2 * sizeof(SheetTotals)) == hipSuccess &&
       hipMalloc(&p->totals_save, 2 * sizeof(SheetTotals)) == hipSuccess &&
       hipMalloc(&p->stats_save, 2 * sizeof(SheetStats)) == hipSuccess &&
       hipEventCreateWithFlags(&p->ev_stats, hipEventDisableTiming) == hipSuccess &&
       hipHostMalloc(&p->h_stats, 2 * sizeof(SheetStats), hipHostMallocDefault) == hipSuccess &&
       hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) == hipSuccess &&
       hipEventCreate(&p->ev0) == hipSuccess && hipEventCreate(&p->ev1) == hipSuccess;
  if (ok) ok = hipMemset(p->totals, 0, 2 * sizeof(SheetTotals)) == hipSuccess;
  for (int l = 0; l < 2 && ok; l++)
    for (int b = 0; b < 2 && ok; b++)
      for (int k = 0; k < 3 && ok; k++) ok = hipMemset(p->planes[l][b][k], 0, fb) == hipSuccess;
  if (!ok) {
    silt_pair_destroy(p);
    return nullptr;
  }
  stats_reset(p->stats, 2, p->stream);
  (void)hipStreamSynchronize(p->stream);
  p->origin[0] = -0.5f * (float)res * cell;
  p->origin[1] = -0.5f * (float)res * cell;
  return p;
}

void silt_pair_destroy(silt_pair* p) {
  if (!p) return;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  plan_labeller_free(p->lab);
  for (PlanScratch& q : p->ps) plan_scratch_free(q);
  if (p->net_ready) mlp_free(p->net);
  for (void* q : {(void*)p->soil_bed, (void*)p->soil_placed, (void*)p->soil_init, (void*)p->soil_rut,
                  (void*)p->d_contacts, (void*)p->sur_ctx, (void*)p->sur_fac, p->d_tires})
    if (q) (void)hipFree(q);
  if (p->h_tires) (void)hipHostFree(p->h_tires);
  for (hipEvent_t e : p->tev)
    if (e) (void)hipEventDestroy(e);
  (void)hipFree(p->bed);
  (void)hipFree(p->nu_slurry);
  for (int l = 0; l < 2; l++)
    for (int b = 0; b < 2; b++)
      for (int k = 0; k < 3; k++) (void)hipFree(p->planes[l][b][k]);
  (void)hipFree(p->stats);
  (void)hipFree(p->totals);
  if (p->d_clocks) (void)hipFree(p->d_clocks);
  if (p->totals_save) (void)hipFree(p->totals_save);
  if (p->stats_save) (void)hipFree(p->stats_save);
  if (p->ev_stats) (void)hipEventDestroy(p->ev_stats);
  if (p->h_stats) (void)hipHostFree(p->h_stats);
  if (p->ev0) (void)hipEventDestroy(p->ev0);
  if (p->ev1) (void)hipEventDestroy(p->ev1);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

int silt_pair_upload(silt_pair* p, const float* bed, const float* water, const float* uw, const float* vw,
                     const float* slurry, const float* us, const float* vs, const float* nu_slurry,
                     float nu_slurry_uniform) {
  if (!p || !bed || !water || !uw || !vw || !slurry || !us || !vs) return SILT_EINVAL;
  bool uniform = isfinite(nu_slurry_uniform);
  if (!uniform && !nu_slurry) return SILT_EINVAL;
  size_t fb = p->n * sizeof(float);
  hipStream_t st = p->stream;
  HIPCHK(hipMemcpyAsync(p->bed, bed, fb, hipMemcpyHostToDevice, st));
  const float* src[2][3] = {{water, uw, vw}, {slurry, us, vs}};
  for (int l = 0; l < 2; l++) {
    p->cur[l] = 0;
    for (int k = 0; k < 3; k++) HIPCHK(hipMemcpyAsync(p->planes[l][0][k], src[l][k], fb, hipMemcpyHostToDevice, st));
  }
  p->nu_slurry_uniform = uniform;
  p->nu_slurry0 = uniform ? nu_slurry_uniform : 0.0f;
  if (!uniform) HIPCHK(hipMemcpyAsync(p->nu_slurry, nu_slurry, fb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(p->totals, 0, 2 * sizeof(SheetTotals), st));
  for (int l = 0; l < 2; l++)
    stats_compute(p->planes[l][0][0], p->planes[l][0][1], p->planes[l][0][2], p->res, kEps, p->stats + l, st);
  HIPCHK(hipMemcpyAsync(p->h_stats, p->stats, 2 * sizeof(SheetStats), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int l = 0; l < 2; l++) p->vol0[l] = p->h_stats[l].vol;
  p->stats_stale = false;
  p->spec_valid = false;
  return SILT_OK;
}

int silt_pair_download(silt_pair* p, float* water, float* uw, float* vw, float* slurry, float* us, float* vs) {
  if (!p || !water || !uw || !vw || !slurry || !us || !vs) return SILT_EINVAL;
  size_t fb = p->n * sizeof(float);
  float* dst[2][3] = {{water, uw, vw}, {slurry, us, vs}};
  for (int l = 0; l < 2; l++)
    for (int k = 0; k < 3; k++)
      HIPCHK(hipMemcpyAsync(dst[l][k], p->planes[l][p->cur[l]][k], fb, hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return SILT_OK;
}

namespace {

// Statistics of the current planes again, when something other than a step
// kernel changed them.
void refresh_stats(silt_pair* p) {
  if (!p->stats_stale) return;
  for (int l = 0; l < 2; l++) {
    float** pl = p->planes[l][p->cur[l]];
    stats_compute(pl[0], pl[1], pl[2], p->res, kEps, p->stats + l, p->stream);
  }
  p->stats_stale = false;
}

}  // namespace

int silt_pair_step(silt_pair* p, float dt) {
  if (!p) return SILT_EINVAL;
  if (!(dt > 0.0f)) return 0;
  hipStream_t st = p->stream;
  refresh_stats(p);
  // Speculation: the last step's plan is launched at once, while the
  // statistics of the current planes come back and the true plan is made;
  // a plan that differs undoes the launch (the current planes are untouched,
  // the ledger lines are restored) and launches the true plan.  The GPU
  // then does not wait for the host's planning between steps.
  bool spec = p->speculate && p->spec_valid && p->spec_dt == dt;
  if (spec && p->spec_pause > 0) {
    p->spec_pause--;
    spec = false;
  }
  int cur0[2] = {p->cur[0], p->cur[1]};
  if (spec) {
    HIPCHK(hipMemcpyAsync(p->h_stats, p->stats, 2 * sizeof(SheetStats), hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(p->ev_stats, st));
    HIPCHK(hipMemcpyAsync(p->totals_save, p->totals, 2 * sizeof(SheetTotals), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(p->stats_save, p->stats, 2 * sizeof(SheetStats), hipMemcpyDeviceToDevice, st));
    int rc = pair_launch(p, dt, p->spec_plans);
    if (rc < 0) return rc;
    HIPCHK(hipEventSynchronize(p->ev_stats));
  } else {
    HIPCHK(hipMemcpyAsync(p->h_stats, p->stats, 2 * sizeof(SheetStats), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  // Plan from the statistics of the planes the step started from.
  int cur1[2] = {p->cur[0], p->cur[1]};
  p->cur[0] = cur0[0];
  p->cur[1] = cur0[1];
  StepPlan plans[2];
  int rc = pair_plan(p, dt, plans);
  p->cur[0] = cur1[0];
  p->cur[1] = cur1[1];
  if (rc < 0) return rc;
  if (spec) {
    if (same_plans(plans, p->spec_plans)) {
      p->spec_hits++;
      p->spec_pause_next = 1;
      return std::max(p->last_nsub[0], p->last_nsub[1]);
    }
    // Undo: back to the planes the step started from, ledger lines and
    // statistics as before (a layer the true plan leaves dry keeps them).
    p->cur[0] = cur0[0];
    p->cur[1] = cur0[1];
    HIPCHK(hipMemcpyAsync(p->totals, p->totals_save, 2 * sizeof(SheetTotals), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(p->stats, p->stats_save, 2 * sizeof(SheetStats), hipMemcpyDeviceToDevice, st));
    p->spec_misses++;
    p->spec_pause = p->spec_pause_next;
    p->spec_pause_next = std::min(64, 2 * p->spec_pause_next);
  }
  rc = pair_launch(p, dt, plans);
  if (rc < 0) return rc;
  p->spec_valid = pair_speculable(p, plans);
  if (p->spec_valid) {
    p->spec_plans[0] = plans[0];
    p->spec_plans[1] = plans[1];
    p->spec_dt = dt;
  }
  return rc;
}

int silt_pair_ledger(silt_pair* p, silt_pair_ledger_t* out) {
  if (!p || !out) return SILT_EINVAL;
  SheetTotals t[2];
  HIPCHK(hipMemcpyAsync(p->h_stats, p->stats, 2 * sizeof(SheetStats), hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipMemcpyAsync(t, p->totals, sizeof t, hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  double area = (double)p->cell * (double)p->cell;
  for (int l = 0; l < 2; l++) {
    silt_layer_ledger_t& L = l == 0 ? out->water : out->slurry;
    L.volume_m3 = p->h_stats[l].vol * area;
    L.initial_m3 = p->vol0[l] * area;
    L.cut_m3 = t[l].cut * area;
    L.residue_m3 = t[l].residue * area;
    L.gap_m3 = (p->h_stats[l].vol + t[l].cut - t[l].residue - p->vol0[l]) * area;
  }
  out->total_m3 = out->water.volume_m3 + out->slurry.volume_m3;
  return SILT_OK;
}

float silt_pair_last_kernel_ms(const silt_pair* p) {
  if (!p) return 0.0f;
  silt_pair* q = const_cast<silt_pair*>(p);
  if (q->kernel_ms_pending) {
    // The step returns while its launch may still run: wait for it here.
    if (hipEventSynchronize(q->ev1) != hipSuccess ||
        hipEventElapsedTime(&q->last_kernel_ms, q->ev0, q->ev1) != hipSuccess)
      q->last_kernel_ms = 0.0f;
    q->kernel_ms_pending = false;
  }
  return q->last_kernel_ms;
}

int silt_pair_set_speculation(silt_pair* p, int on) {
  if (!p) return SILT_EINVAL;
  p->speculate = on != 0;
  p->spec_valid = false;
  p->spec_pause = 0;
  p->spec_pause_next = 1;
  return SILT_OK;
}

int silt_pair_speculation_counts(const silt_pair* p, long long* hits, long long* misses) {
  if (!p) return SILT_EINVAL;
  if (hits) *hits = p->spec_hits;
  if (misses) *misses = p->spec_misses;
  return SILT_OK;
}

// The caller's state around tuning steps: the current planes, statistics,
// ledger lines and speculation state, put back by restore().
struct PairSnapshot {
  silt_pair* p = nullptr;
  float* save = nullptr;
  int cur0[2] = {0, 0};
  SheetStats stats0[2];
  SheetTotals totals0[2];
  long long hits0 = 0, misses0 = 0;
  int pause0 = 0, pause_next0 = 1;

  int take(silt_pair* q) {
    p = q;
    const size_t fb = p->n * sizeof(float);
    if (hipMalloc(&save, 6 * fb) != hipSuccess) {
      save = nullptr;
      return SILT_ENOMEM;
    }
    hipStream_t st = p->stream;
    cur0[0] = p->cur[0];
    cur0[1] = p->cur[1];
    refresh_stats(p);
    hits0 = p->spec_hits;
    misses0 = p->spec_misses;
    pause0 = p->spec_pause;
    pause_next0 = p->spec_pause_next;
    for (int l = 0; l < 2; l++)
      for (int k = 0; k < 3; k++)
        HIPCHK(hipMemcpyAsync(save + (size_t)(3 * l + k) * p->n, p->planes[l][cur0[l]][k], fb,
                              hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(stats0, p->stats, sizeof stats0, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(totals0, p->totals, sizeof totals0, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
  }
  int restore() {
    const size_t fb = p->n * sizeof(float);
    hipStream_t st = p->stream;
    p->cur[0] = cur0[0];
    p->cur[1] = cur0[1];
    p->spec_valid = false;
    p->spec_hits = hits0;
    p->spec_misses = misses0;
    p->spec_pause = pause0;
    p->spec_pause_next = pause_next0;
    for (int l = 0; l < 2; l++)
      for (int k = 0; k < 3; k++)
        HIPCHK(hipMemcpyAsync(p->planes[l][cur0[l]][k], save + (size_t)(3 * l + k) * p->n, fb,
                              hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(p->stats, stats0, sizeof stats0, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(p->totals, totals0, sizeof totals0, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
  }
  ~PairSnapshot() {
    if (save) (void)hipFree(save);
  }
};

int silt_pair_autotune(silt_pair* p, float dt, const int* rows, size_t n, int reps, int* best_rows, float* best_ms) {
  if (!p || !(dt > 0.0f) || reps < 1 || reps > 1000 || (n && !rows)) return SILT_EINVAL;
  std::vector<int> cand;
  if (rows && n) {
    for (size_t k = 0; k < n; k++)
      if (rows[k] >= 0) cand.push_back(rows[k]);
  } else {
    // The default (one round of resident tiles) and heights around it.
    cand = {0, 48, 64, 96, 128, 160, 192, 256, 384, 512};
  }
  if (cand.empty()) return SILT_EINVAL;
  // The tuning steps are not the caller's: their planes, speculation counts
  // and plans are dropped.
  PairSnapshot snap;
  int rc = snap.take(p);
  if (rc != 0) return rc;
  // Heights are tuned for the joint launch of both layers.
  const int pinned0 = fused_pinned_chunk_rows(p->res, 2);
  float best = INFINITY;
  int best_c = pinned0;
  for (size_t c = 0; c < cand.size() && rc == 0; c++) {
    fused_set_chunk_rows_for(p->res, cand[c], 2);
    rc = silt_pair_step(p, dt);  // warm-up of this height
    double sum = 0.0;
    for (int k = 0; k < reps && rc >= 0; k++) {
      rc = silt_pair_step(p, dt);
      sum += silt_pair_last_kernel_ms(p);
    }
    int rr = snap.restore();
    if (rc >= 0) rc = rr;
    if (rc < 0) break;
    rc = 0;
    float mean = (float)(sum / reps);
    if (mean < best) {
      best = mean;
      best_c = cand[c];
    }
  }
  fused_set_chunk_rows_for(p->res, rc == 0 ? best_c : pinned0, 2);
  if (rc < 0) return rc;
  if (best_rows) *best_rows = best_c;
  if (best_ms) *best_ms = best;
  return SILT_OK;
}

int silt_pair_set_xcd_weights(silt_pair* p, const float* w) {
  if (!p) return SILT_EINVAL;
  if (!w) {
    p->xcd_on = false;
    for (int x = 0; x < kXcds; x++) p->xcd_w[x] = 1.0f;
    return SILT_OK;
  }
  for (int x = 0; x < kXcds; x++)
    if (!(w[x] >= 0.25f && w[x] <= 4.0f)) return SILT_EINVAL;
  for (int x = 0; x < kXcds; x++) p->xcd_w[x] = w[x];
  p->xcd_on = true;
  return SILT_OK;
}

namespace {

// One calibration round: reps steps with the current weights, each
// recording its workgroups' clocks.  Per XCD: when its last workgroup ended
// (microseconds after the launch's first start: the XCD's finish), the
// mean pipeline rows of its tiles and the mean tile time; and the mean
// kernel time of the steps.
struct XcdRound {
  double finish[kXcds] = {0}, rows[kXcds] = {0}, tile_us[kXcds] = {0};
  double kernel_ms = 0.0;
};

int xcd_round(silt_pair* p, float dt, int reps, size_t nrec, XcdRound* out) {
  std::vector<unsigned long long>& rec = p->h_clocks;
  rec.assign(4 * nrec, 0ull);
  XcdRound r;
  double cnt[kXcds] = {0};
  p->clocks_on = true;
  int rc = silt_pair_step(p, dt);  // first launch of these weights, not counted
  for (int k = 0; k < reps && rc >= 0; k++) {
    if (hipMemsetAsync(p->d_clocks, 0, rec.size() * sizeof(unsigned long long), p->stream) != hipSuccess)
      rc = SILT_EDEVICE;
    if (rc >= 0) rc = silt_pair_step(p, dt);
    if (rc >= 0) r.kernel_ms += silt_pair_last_kernel_ms(p) / reps;
    if (rc >= 0 && (hipMemcpyAsync(rec.data(), p->d_clocks, rec.size() * sizeof(unsigned long long),
                                   hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
                    hipStreamSynchronize(p->stream) != hipSuccess))
      rc = SILT_EDEVICE;
    if (rc < 0) break;
    unsigned long long t0 = ~0ull;
    for (size_t g = 0; g < nrec; g++)
      if (rec[4 * g + 1] > rec[4 * g]) t0 = std::min(t0, rec[4 * g]);
    double fin[kXcds] = {0};
    // Workgroup g ran on XCD g mod 8 (round-robin dispatch).
    for (size_t g = 0; g < nrec; g++) {
      const unsigned long long* c = &rec[4 * g];
      if (c[1] <= c[0]) continue;
      const int x = (int)(g % kXcds);
      fin[x] = std::max(fin[x], (double)(c[1] - t0) * 0.01);  // 100 MHz ticks -> us
      r.tile_us[x] += (double)(c[1] - c[0]) * 0.01;
      r.rows[x] += (double)(c[2] & 0xffffffffull);
      cnt[x] += 1.0;
    }
    for (int x = 0; x < kXcds; x++) r.finish[x] += fin[x] / reps;
  }
  p->clocks_on = false;
  if (rc < 0) return rc;
  for (int x = 0; x < kXcds; x++) {
    if (cnt[x] == 0.0 || r.finish[x] <= 0.0) return SILT_EUNSUPPORTED;
    r.tile_us[x] /= cnt[x];
    r.rows[x] /= cnt[x];
  }
  *out = r;
  return 0;
}

}  // namespace

size_t silt_pair_clock_records(const silt_pair* p, unsigned long long* out, size_t n) {
  if (!p) return 0;
  const size_t have = p->h_clocks.size() / 4;
  if (out) memcpy(out, p->h_clocks.data(), std::min(n, have) * 4 * sizeof(unsigned long long));
  return have;
}

int silt_pair_balance(silt_pair* p, float dt, int reps, float* weights, float* ms, float* tile_us) {
  if (!p || !(dt > 0.0f) || reps < 1 || reps > 1000) return SILT_EINVAL;
  if (p->path == SILT_PATH_EXACT) return SILT_EUNSUPPORTED;
  // Record room for the most tiles a launch of this side can have (two
  // layers, chunks of at least 8 rows).
  const int strips = (p->res + 234) / 235;
  const size_t nrec = (size_t)2 * strips * ((p->res + 7) / 8);
  if (!p->d_clocks && hipMalloc(&p->d_clocks, 4 * nrec * sizeof(unsigned long long)) != hipSuccess) {
    p->d_clocks = nullptr;
    return SILT_ENOMEM;
  }
  PairSnapshot snap;
  int rc = snap.take(p);
  if (rc != 0) return rc;
  const bool on0 = p->xcd_on;
  float w0[kXcds];
  memcpy(w0, p->xcd_w, sizeof w0);
  // Round 0 runs uniform chunks.  Each later round moves rows toward the
  // XCDs that finished early: with u[x] the mean pipeline rows of XCD x's
  // tiles (output rows n[x] plus the fill F around them) and f[x] when its
  // last workgroup ended, u'[x] is proportional to u[x] f[x]^-a with the
  // total rows kept, and the weights are n'[x] = u'[x] - F relative to
  // their mean (an XCD's rate moves with its rows, so this is a search,
  // not a solve).  Every
  // round is scored by the launch's mean kernel time and the fastest
  // weights are kept, uniform chunks included.
  const double fill = 2.0 * 15.0;
  float w[kXcds] = {1, 1, 1, 1, 1, 1, 1, 1}, best_w[kXcds] = {1, 1, 1, 1, 1, 1, 1, 1};
  double best_ms = INFINITY, t_uni = 0.0;
  XcdRound first, last, best_r;
  p->xcd_on = true;
  // A hill climb: each round's weights come from the best round so far,
  // with the step (the power of the finish-time ratio) halved after a round
  // that did not beat it.
  const int rounds = 12;
  double alpha = 1.0;
  for (int round = 0; round < rounds && rc >= 0; round++) {
    memcpy(p->xcd_w, w, sizeof w);
    XcdRound r;
    rc = xcd_round(p, dt, reps, nrec, &r);
    if (rc < 0) break;
    rc = snap.restore();
    if (rc < 0) break;
    if (round == 0) {
      first = r;
      t_uni = r.kernel_ms;
    }
    last = r;
    if (r.kernel_ms < best_ms) {
      best_ms = r.kernel_ms;
      memcpy(best_w, w, sizeof w);
      best_r = r;
    } else {
      alpha *= 0.5;
    }
    double rows = 0.0, scaled = 0.0, u[kXcds];
    for (int x = 0; x < kXcds; x++) {
      u[x] = best_r.rows[x] * std::pow(best_r.finish[x], -alpha);
      rows += best_r.rows[x];
      scaled += u[x];
    }
    double n[kXcds], nsum = 0.0;
    for (int x = 0; x < kXcds; x++) {
      n[x] = std::max(8.0, u[x] * rows / scaled - fill);
      nsum += n[x];
    }
    double wsum = 0.0;
    for (int x = 0; x < kXcds; x++) {
      w[x] = (float)std::min(2.5, std::max(0.4, n[x] * kXcds / nsum));
      wsum += w[x];
    }
    for (int x = 0; x < kXcds; x++) w[x] = (float)(w[x] * kXcds / wsum);
  }
  if (rc < 0) {
    p->xcd_on = on0;
    memcpy(p->xcd_w, w0, sizeof w0);
    return rc;
  }
  const bool keep = std::any_of(best_w, best_w + kXcds, [](float v) { return v != 1.0f; });
  p->xcd_on = keep;
  memcpy(p->xcd_w, best_w, sizeof best_w);
  if (weights) memcpy(weights, p->xcd_w, sizeof p->xcd_w);
  if (ms) {
    ms[0] = (float)t_uni;
    ms[1] = (float)best_ms;
  }
  if (tile_us) {
    for (int x = 0; x < kXcds; x++) tile_us[x] = (float)first.finish[x];
    for (int x = 0; x < kXcds; x++) tile_us[kXcds + x] = (float)last.finish[x];
  }
  return SILT_OK;
}

int silt_pair_last_fallback(const silt_pair* p) { return p ? p->last_fallback : SILT_EINVAL; }

int silt_pair_set_ground(silt_pair* p, const float* bed, float origin_x, float origin_z) {
  if (!p || !bed || !isfinite(origin_x) || !isfinite(origin_z)) return SILT_EINVAL;
  const size_t fb = p->n * sizeof(float);
  if (!p->soil_bed) {
    if (hipMalloc(&p->soil_bed, fb) != hipSuccess || hipMalloc(&p->soil_placed, fb) != hipSuccess ||
        hipMalloc(&p->soil_init, fb) != hipSuccess || hipMalloc(&p->soil_rut, fb) != hipSuccess)
      return SILT_ENOMEM;
  }
  HIPCHK(hipMemcpyAsync(p->soil_bed, bed, fb, hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipMemcpyAsync(p->soil_init, bed, fb, hipMemcpyHostToDevice, p->stream));
  HIPCHK(hipMemsetAsync(p->soil_placed, 0, fb, p->stream));
  core_ground(p->soil_bed, p->soil_placed, p->soil_init, p->bed, p->soil_rut, (int)p->n, p->stream);
  HIPCHK(hipStreamSynchronize(p->stream));
  p->origin[0] = origin_x;
  p->origin[1] = origin_z;
  return SILT_OK;
}

int silt_pair_get_ground(silt_pair* p, float* bed, float* placed, float* rut) {
  if (!p || !p->soil_bed) return SILT_EINVAL;
  const size_t fb = p->n * sizeof(float);
  if (bed) HIPCHK(hipMemcpyAsync(bed, p->soil_bed, fb, hipMemcpyDeviceToHost, p->stream));
  if (placed) HIPCHK(hipMemcpyAsync(placed, p->soil_placed, fb, hipMemcpyDeviceToHost, p->stream));
  if (rut) HIPCHK(hipMemcpyAsync(rut, p->soil_rut, fb, hipMemcpyDeviceToHost, p->stream));
  HIPCHK(hipStreamSynchronize(p->stream));
  return SILT_OK;
}

int silt_pair_set_surrogate(silt_pair* p, const float* params) {
  if (!p) return SILT_EINVAL;
  if (!params) {
    p->net_ready = false;
    return SILT_OK;
  }
  if (!p->tev[0])
    for (hipEvent_t& e : p->tev) HIPCHK(hipEventCreate(&e));
  if (!p->net.master && mlp_init(p->net, (size_t)kSurMax * kSurMax, 0u, p->stream) != 0) return SILT_ENOMEM;
  HIPCHK(hipMemcpyAsync(p->net.master, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, p->stream));
  mlp_sync_bf16(p->net);
  HIPCHK(hipStreamSynchronize(p->stream));
  p->net_ready = true;
  return SILT_OK;
}

float silt_pair_last_tire_ms(const silt_pair* p) { return p ? p->last_tire_ms : 0.0f; }

int silt_pair_set_tire_batching(silt_pair* p, int on) {
  if (!p) return SILT_EINVAL;
  p->tire_batching = on != 0;
  return SILT_OK;
}

int silt_pair_step_tires(silt_pair* p, float dt, const silt_tire_push* pushes, size_t npush, size_t per_vehicle,
                         const silt_tire_contact* contacts, size_t ncontacts) {
  if (!p || (npush && (!pushes || per_vehicle == 0 || npush % per_vehicle || per_vehicle > (size_t)kMaxPush)) ||
      (ncontacts && !contacts))
    return SILT_EINVAL;
  if (!(dt > 0.0f)) return 0;
  hipStream_t st = p->stream;
  const int res = p->res;
  // Ruts first: the contacts cut the ground, the sheet bed follows.
  if (ncontacts && p->soil_bed) {
    if (p->cap_contacts < ncontacts) {
      HIPCHK(hipStreamSynchronize(st));
      if (p->d_contacts) (void)hipFree(p->d_contacts);
      p->d_contacts = nullptr;
      p->cap_contacts = 0;
      if (hipMalloc(&p->d_contacts, ncontacts * sizeof(RutContact)) != hipSuccess) return SILT_ENOMEM;
      p->cap_contacts = ncontacts;
    }
    // Contacts come in the pushes' vehicle groups; vehicles' footprints do
    // not overlap, so each vehicle stamps in its own workgroup.
    const int per = (npush && ncontacts == npush) ? (int)per_vehicle : 0;
    // The stamp refreshes the sheet bed and the rut over its footprints.
    int rc = rut_stamp(p->soil_bed, p->soil_placed, p->soil_init, res, p->cell, p->origin[0], p->origin[1], contacts,
                       ncontacts, dt, p->d_contacts, st, per, p->bed, p->soil_rut);
    if (rc) return rc;
  }
  // Vehicles: groups of per_vehicle pushes.
  const size_t groups = npush ? npush / per_vehicle : 0;
  std::vector<PushList> lists(groups);
  std::vector<TireJob> jobs(groups);
  std::vector<CorePush> flat(npush);
  size_t cells = 0;
  int max_w = 0, max_h = 0;
  for (size_t v = 0; v < groups; v++) {
    PushList& pl = lists[v];
    pl.n = (int)per_vehicle;
    for (size_t k = 0; k < per_vehicle; k++) {
      const silt_tire_push& q = pushes[v * per_vehicle + k];
      pl.p[k] = CorePush{q.x, q.z, q.forward_x, q.forward_z, q.radius, q.width, q.speed, q.slip};
      flat[v * per_vehicle + k] = pl.p[k];
    }
    TireJob& J = jobs[v];
    J.r = p->net_ready ? sur_region_at(pl.p, pl.n, p->origin[0], p->origin[1], p->cell, res) : SurRegion{};
    J.row = (int)cells;
    cells += (size_t)J.r.nx * J.r.nz;
    sur_mean_velocity(pl.p, pl.n, &J.vtx, &J.vtz);
    // The cells push_on (src/fluid_core.rs:8152-8239) can reach around the
    // vehicle's pushes.
    float x0 = 1e30f, x1 = -1e30f, z0 = 1e30f, z1 = -1e30f;
    for (int k = 0; k < pl.n; k++) {
      const CorePush& q = pl.p[k];
      const float reach = q.radius + q.width + 0.5f;
      x0 = fminf(x0, q.x - reach);
      x1 = fmaxf(x1, q.x + reach);
      z0 = fminf(z0, q.z - reach);
      z1 = fmaxf(z1, q.z + reach);
    }
    int i0 = std::max(0, (int)floorf((x0 - p->origin[0]) / p->cell));
    int j0 = std::max(0, (int)floorf((z0 - p->origin[1]) / p->cell));
    int i1 = std::min(res, (int)ceilf((x1 - p->origin[0]) / p->cell) + 1);
    int j1 = std::min(res, (int)ceilf((z1 - p->origin[1]) / p->cell) + 1);
    J.push = PushBox{(int)(v * per_vehicle), pl.n, i0, j0, std::max(i1 - i0, 0), std::max(j1 - j0, 0)};
    max_w = std::max(max_w, J.push.w);
    max_h = std::max(max_h, J.push.h);
  }
  // Vehicles whose network regions and push boxes are pairwise disjoint run
  // every stage in one launch; otherwise one vehicle after another, in order.
  auto apart = [](int a0, int an, int b0, int bn) { return a0 + an <= b0 || b0 + bn <= a0; };
  bool batched = p->tire_batching && groups > 1;
  for (size_t a = 0; a < groups && batched; a++)
    for (size_t b = a + 1; b < groups && batched; b++) {
      const TireJob &A = jobs[a], &B = jobs[b];
      const bool r_ok = A.r.nx == 0 || B.r.nx == 0 || apart(A.r.i0, A.r.nx, B.r.i0, B.r.nx) ||
                        apart(A.r.j0, A.r.nz, B.r.j0, B.r.nz);
      const bool p_ok = A.push.w == 0 || A.push.h == 0 || B.push.w == 0 || B.push.h == 0 ||
                        apart(A.push.i0, A.push.w, B.push.i0, B.push.w) ||
                        apart(A.push.j0, A.push.h, B.push.j0, B.push.h);
      batched = r_ok && p_ok;
    }
  const TireJob* d_jobs = nullptr;
  const PushBox* d_boxes = nullptr;
  const CorePush* d_push = nullptr;
  if (batched) {
    // Table: the jobs, their push boxes, the pushes (256-byte aligned parts).
    const size_t jb = (groups * sizeof(TireJob) + 255) & ~(size_t)255;
    const size_t bb = (groups * sizeof(PushBox) + 255) & ~(size_t)255;
    const size_t bytes = jb + bb + npush * sizeof(CorePush);
    if (p->cap_tires < bytes) {
      HIPCHK(hipStreamSynchronize(st));
      if (p->d_tires) (void)hipFree(p->d_tires);
      if (p->h_tires) (void)hipHostFree(p->h_tires);
      p->d_tires = p->h_tires = nullptr;
      p->cap_tires = 0;
      if (hipMalloc(&p->d_tires, bytes) != hipSuccess ||
          hipHostMalloc(&p->h_tires, bytes, hipHostMallocDefault) != hipSuccess)
        return SILT_ENOMEM;
      p->cap_tires = bytes;
    }
    // The previous call's copy has completed: every step ends with a wait.
    char* h = static_cast<char*>(p->h_tires);
    memcpy(h, jobs.data(), groups * sizeof(TireJob));
    for (size_t v = 0; v < groups; v++) reinterpret_cast<PushBox*>(h + jb)[v] = jobs[v].push;
    memcpy(h + jb + bb, flat.data(), npush * sizeof(CorePush));
    HIPCHK(hipMemcpyAsync(p->d_tires, p->h_tires, bytes, hipMemcpyHostToDevice, st));
    const char* d = static_cast<const char*>(p->d_tires);
    d_jobs = reinterpret_cast<const TireJob*>(d);
    d_boxes = reinterpret_cast<const PushBox*>(d + jb);
    d_push = reinterpret_cast<const CorePush*>(d + jb + bb);
  }
  const bool timed = p->tev[0] != nullptr;
  if (timed) HIPCHK(hipEventRecord(p->tev[0], st));
  float** wl = p->planes[0][p->cur[0]];
  float** sl = p->planes[1][p->cur[1]];
  if (p->net_ready && cells) {
    if (p->cap_groups < groups) {
      HIPCHK(hipStreamSynchronize(st));
      if (p->sur_ctx) (void)hipFree(p->sur_ctx);
      if (p->sur_fac) (void)hipFree(p->sur_fac);
      p->sur_ctx = p->sur_fac = nullptr;
      p->cap_groups = 0;
      if (hipMalloc(&p->sur_ctx, groups * kSurCtxFloats * sizeof(float)) != hipSuccess ||
          hipMalloc(&p->sur_fac, groups * 2 * kSurMax * kSurMax * sizeof(float)) != hipSuccess)
        return SILT_ENOMEM;
      p->cap_groups = groups;
    }
    int rc = mlp_reserve(p->net, cells);
    if (rc) return rc;
    const SurGrid g{wl[0], sl[0], wl[1], wl[2], p->bed, res, p->cell, p->origin[0], p->origin[1]};
    if (batched) {
      sur_context_many(g, d_jobs, d_push, (int)groups, p->sur_ctx, st);
      sur_gather_many(p->sur_ctx, d_jobs, (int)groups, p->net.xb, st);
    } else {
      for (size_t v = 0; v < groups; v++) {
        if (!jobs[v].r.nx) continue;
        float* ctx = p->sur_ctx + v * kSurCtxFloats;
        sur_context(g, jobs[v].r, lists[v], ctx, st);
        sur_gather_region(ctx, jobs[v].r, p->net.xb + (size_t)jobs[v].row * kMlpIn, st);
      }
    }
    mlp_infer(p->net, cells);
  }
  // The pushes shove both layers around each vehicle (push_on, src/fluid_core.rs:8152-8239).
  const GridGeo geo{res, p->cell, p->origin[0], p->origin[1]};
  if (batched) {
    core_push_on_many(geo, wl[0], sl[0], wl[1], wl[2], sl[1], sl[2], d_boxes, d_push, (int)groups, max_w, max_h, st);
  } else {
    for (size_t v = 0; v < groups; v++) {
      const PushBox& b = jobs[v].push;
      if (b.w > 0 && b.h > 0)
        core_push_on(geo, wl[0], sl[0], wl[1], wl[2], sl[1], sl[2], lists[v], b.i0, b.j0, b.w, b.h, st);
    }
  }
  if (timed) HIPCHK(hipEventRecord(p->tev[1], st));
  // Fresh statistics for the plan (the shove changed speeds, the last
  // fluxes depths), then the step.
  if (groups) {
    stats_compute(wl[0], wl[1], wl[2], res, kEps, p->stats, st);
    stats_compute(sl[0], sl[1], sl[2], res, kEps, p->stats + 1, st);
    p->stats_stale = false;
  }
  int nsub = silt_pair_step(p, dt);
  if (nsub < 0) return nsub;
  if (timed) HIPCHK(hipEventRecord(p->tev[2], st));
  if (p->net_ready && cells) {
    float* wh = p->planes[0][p->cur[0]][0];
    float* sh = p->planes[1][p->cur[1]][0];
    if (batched) {
      sur_apply_many(wh, sh, res, p->net.y, d_jobs, (int)groups, p->sur_fac, st, p->totals, p->stats);
    } else {
      for (size_t v = 0; v < groups; v++) {
        if (!jobs[v].r.nx) continue;
        sur_apply(wh, sh, res, p->net.y + (size_t)jobs[v].row * kMlpOut, jobs[v].r,
                  p->sur_fac + v * 2 * kSurMax * kSurMax, st, p->totals, p->stats);
      }
    }
    p->stats_stale = true;
  }
  if (timed) {
    HIPCHK(hipEventRecord(p->tev[3], st));
    HIPCHK(hipEventSynchronize(p->tev[3]));
    float a = 0.0f, b = 0.0f;
    HIPCHK(hipEventElapsedTime(&a, p->tev[0], p->tev[1]));
    HIPCHK(hipEventElapsedTime(&b, p->tev[2], p->tev[3]));
    p->last_tire_ms = a + b;
  }
  return nsub;
}

}  // extern "C"
