// Batch of independent sheet windows (include/silt_batch.h).  Every window
// is one layer stepped as the reference's generic-res step_sheet
// (src/fluid_core.rs:770-785, Gravity drive).  A step plans each window from
// the statistics its previous kernel left (the wet box and largest speed),
// then runs substep s of every window that has one in a single launch of the
// fused kernel.  The step's launches (a statistics reset, the descriptor
// copies from pinned staging and the kernels) are captured once in a
// hipGraph and replayed while their shape (windows per launch, tiles, rows
// per workgroup, kernel instance) is unchanged; only the staged descriptors
// change between replays.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "plan.hpp"
#include "sheet_device.hpp"
#include "sheet_fused.hpp"
#include "silt_batch.h"
#include "silt_sheet.h"
#include "workspace.hpp"

#define HIPCHK(x)                                 \
  do {                                            \
    if ((x) != hipSuccess) return SILT_EDEVICE;   \
  } while (0)

using namespace silt;

namespace {

const float kEps = 1.0e-6f;
constexpr int kMaxLaunches = 8;  // substeps per step (nsub cap)

struct Window {
  int res = 0;
  float cell = 0.2f;
  size_t n = 0;
  float* bed = nullptr;
  float* planes[2][3] = {};  // [buffer][h, u, v]
  int cur = 0;
  float nu = 0.9f;
  float vort = 0.08f;
  double vol0 = 0.0;
};

// Device copy and pinned staging of one launch's descriptors.
struct Slot {
  unsigned char* d_table = nullptr;
  int* d_tiles = nullptr;
  int* d_rows = nullptr;  // [2 tiles] rows of every tile under XCD weights
  unsigned char* h_table = nullptr;
  int* h_tiles = nullptr;
  int* h_rows = nullptr;
  size_t cap_table = 0, cap_tiles = 0, cap_rows = 0;
};

}  // namespace

struct silt_batch {
  int path = SILT_PATH_FAST;
  std::vector<Window> w;
  SheetStats* stats = nullptr;    // [n] device
  SheetTotals* totals = nullptr;  // [n] device
  SheetStats* h_stats = nullptr;  // [n] pinned
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Slot slot[kMaxLaunches];
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<long long> graph_key;
  int graphs = 0;
  int rows = 0;
  // XCD weights (silt_batch_set_xcd_weights): xcd_have once given, xcd_on
  // while the launches use them (the autotuner keeps them only when faster).
  float xcd_w[kXcds] = {1, 1, 1, 1, 1, 1, 1, 1};
  bool xcd_have = false, xcd_on = false;
  float last_ms = 0.0f;
  int last_fallbacks = 0;
  PlanLabeller lab;
  std::vector<PlanScratch> ps;  // [n]
};

namespace {

int grow(Slot& s, size_t table_bytes, size_t tiles, size_t rows) {
  if (s.cap_table < table_bytes) {
    if (s.d_table) (void)hipFree(s.d_table);
    if (s.h_table) (void)hipHostFree(s.h_table);
    s.d_table = nullptr;
    s.h_table = nullptr;
    s.cap_table = 0;
    if (hipMalloc(&s.d_table, table_bytes) != hipSuccess ||
        hipHostMalloc(&s.h_table, table_bytes, hipHostMallocDefault) != hipSuccess)
      return SILT_ENOMEM;
    s.cap_table = table_bytes;
  }
  size_t want = tiles ? tiles : 1;
  if (s.cap_tiles < want) {
    if (s.d_tiles) (void)hipFree(s.d_tiles);
    if (s.h_tiles) (void)hipHostFree(s.h_tiles);
    s.d_tiles = nullptr;
    s.h_tiles = nullptr;
    s.cap_tiles = 0;
    if (hipMalloc(&s.d_tiles, want * sizeof(int)) != hipSuccess ||
        hipHostMalloc(&s.h_tiles, want * sizeof(int), hipHostMallocDefault) != hipSuccess)
      return SILT_ENOMEM;
    s.cap_tiles = want;
  }
  if (s.cap_rows < rows) {
    if (s.d_rows) (void)hipFree(s.d_rows);
    if (s.h_rows) (void)hipHostFree(s.h_rows);
    s.d_rows = nullptr;
    s.h_rows = nullptr;
    s.cap_rows = 0;
    if (hipMalloc(&s.d_rows, rows * sizeof(int)) != hipSuccess ||
        hipHostMalloc(&s.h_rows, rows * sizeof(int), hipHostMallocDefault) != hipSuccess)
      return SILT_ENOMEM;
    s.cap_rows = rows;
  }
  return 0;
}

void drop_graph(silt_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_key.clear();
}

// The launches of one step, in order: statistics reset, then per launch the
// descriptor copies and the kernel.
int enqueue_step(silt_batch* b, const std::vector<FusedLaunch>& launches) {
  hipStream_t st = b->stream;
  stats_reset(b->stats, (int)b->w.size(), st);
  for (size_t s = 0; s < launches.size(); s++) {
    const FusedLaunch& l = launches[s];
    Slot& sl = b->slot[s];
    HIPCHK(hipMemcpyAsync(sl.d_table, sl.h_table, l.descs.size(), hipMemcpyHostToDevice, st));
    if (!l.tile_window.empty())
      HIPCHK(hipMemcpyAsync(sl.d_tiles, sl.h_tiles, l.tile_window.size() * sizeof(int), hipMemcpyHostToDevice, st));
    if (!l.rows.empty())
      HIPCHK(hipMemcpyAsync(sl.d_rows, sl.h_rows, l.rows.size() * sizeof(int), hipMemcpyHostToDevice, st));
    int rc = fused_enqueue(l, sl.d_table, sl.d_tiles, st, l.rows.empty() ? nullptr : sl.d_rows);
    if (rc != 0) return rc;
  }
  return 0;
}

}  // namespace

extern "C" {

silt_batch* silt_batch_create(size_t n, const size_t* res, const float* cell, int path) {
  if (n < 1 || n > 4096 || !res || !cell) return nullptr;
  if (path != SILT_PATH_FAST && path != SILT_PATH_FUSED_EXACT) return nullptr;
  for (size_t k = 0; k < n; k++)
    if (res[k] < 8 || res[k] > 16384 || !(cell[k] > kFineCell + 1.0e-6f)) return nullptr;
  silt_batch* b = new silt_batch();
  b->path = path;
  b->w.resize(n);
  b->ps.resize(n);
  bool ok = true;
  for (size_t k = 0; k < n && ok; k++) {
    Window& w = b->w[k];
    w.res = (int)res[k];
    w.cell = cell[k];
    w.n = res[k] * res[k];
    size_t fb = w.n * sizeof(float);
    ok = hipMalloc(&w.bed, fb) == hipSuccess;
    for (int q = 0; q < 2 && ok; q++)
      for (int c = 0; c < 3 && ok; c++)
        ok = hipMalloc(&w.planes[q][c], fb) == hipSuccess && hipMemset(w.planes[q][c], 0, fb) == hipSuccess;
    ok = ok && hipMemset(w.bed, 0, fb) == hipSuccess;
  }
  ok = ok && hipMalloc(&b->stats, n * sizeof(SheetStats)) == hipSuccess &&
       hipMalloc(&b->totals, n * sizeof(SheetTotals)) == hipSuccess &&
       hipHostMalloc(&b->h_stats, n * sizeof(SheetStats), hipHostMallocDefault) == hipSuccess &&
       hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
       hipEventCreate(&b->ev0) == hipSuccess && hipEventCreate(&b->ev1) == hipSuccess &&
       hipMemset(b->totals, 0, n * sizeof(SheetTotals)) == hipSuccess;
  if (!ok) {
    silt_batch_destroy(b);
    return nullptr;
  }
  stats_reset(b->stats, (int)n, b->stream);
  if (hipMemcpyAsync(b->h_stats, b->stats, n * sizeof(SheetStats), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess) {
    silt_batch_destroy(b);
    return nullptr;
  }
  return b;
}

void silt_batch_destroy(silt_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  plan_labeller_free(b->lab);
  for (PlanScratch& p : b->ps) plan_scratch_free(p);
  for (Window& w : b->w) {
    (void)hipFree(w.bed);
    for (int q = 0; q < 2; q++)
      for (int c = 0; c < 3; c++) (void)hipFree(w.planes[q][c]);
  }
  for (Slot& s : b->slot) {
    (void)hipFree(s.d_table);
    (void)hipFree(s.d_tiles);
    if (s.h_table) (void)hipHostFree(s.h_table);
    if (s.h_tiles) (void)hipHostFree(s.h_tiles);
    if (s.d_rows) (void)hipFree(s.d_rows);
    if (s.h_rows) (void)hipHostFree(s.h_rows);
  }
  (void)hipFree(b->stats);
  (void)hipFree(b->totals);
  if (b->h_stats) (void)hipHostFree(b->h_stats);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int silt_batch_upload(silt_batch* b, size_t k, const float* bed, const float* h, const float* u, const float* v,
                      float nu, float vorticity) {
  if (!b || k >= b->w.size() || !bed || !h || !u || !v || !isfinite(nu) || !isfinite(vorticity)) return SILT_EINVAL;
  Window& w = b->w[k];
  size_t fb = w.n * sizeof(float);
  hipStream_t st = b->stream;
  w.cur = 0;
  w.nu = nu;
  w.vort = vorticity;
  HIPCHK(hipMemcpyAsync(w.bed, bed, fb, hipMemcpyHostToDevice, st));
  const float* src[3] = {h, u, v};
  for (int c = 0; c < 3; c++) HIPCHK(hipMemcpyAsync(w.planes[0][c], src[c], fb, hipMemcpyHostToDevice, st));
  SheetTotals zero{0.0, 0.0};
  HIPCHK(hipMemcpyAsync(b->totals + k, &zero, sizeof zero, hipMemcpyHostToDevice, st));
  stats_compute(w.planes[0][0], w.planes[0][1], w.planes[0][2], w.res, kEps, b->stats + k, st);
  HIPCHK(hipMemcpyAsync(b->h_stats + k, b->stats + k, sizeof(SheetStats), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  w.vol0 = b->h_stats[k].vol;
  return SILT_OK;
}

int silt_batch_download(silt_batch* b, size_t k, float* h, float* u, float* v) {
  if (!b || k >= b->w.size() || !h || !u || !v) return SILT_EINVAL;
  Window& w = b->w[k];
  size_t fb = w.n * sizeof(float);
  float* dst[3] = {h, u, v};
  for (int c = 0; c < 3; c++)
    HIPCHK(hipMemcpyAsync(dst[c], w.planes[w.cur][c], fb, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return SILT_OK;
}

int silt_batch_step(silt_batch* b, float dt) {
  if (!b) return SILT_EINVAL;
  if (!(dt > 0.0f)) return 0;
  hipStream_t st = b->stream;
  const size_t n = b->w.size();
  // Plans from the statistics of the current planes (left by the previous
  // step's kernels, or by the upload).
  HIPCHK(hipMemcpyAsync(b->h_stats, b->stats, n * sizeof(SheetStats), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // Every window's boxes come from one batched labelling, and every box
  // maximum is enqueued before the second wait: two waits for the batch.
  std::vector<StepPlan> plans(n);
  std::vector<char> host(n, 0);
  b->last_fallbacks = 0;
  std::vector<PlanSheet> sheets(n);
  for (size_t k = 0; k < n; k++) {
    Window& w = b->w[k];
    sheets[k] = PlanSheet{w.planes[w.cur][0], w.res, w.cell, b->h_stats + k};
  }
  int pr = plan_boxes_many(b->lab, sheets.data(), (int)n, dt, kEps, plans.data(), host.data(), st);
  if (pr < 0) return pr;
  bool wait = false;
  for (size_t k = 0; k < n; k++) {
    Window& w = b->w[k];
    if (host[k]) {
      // Too many wet clusters to read back: the planes go to the host.
      b->last_fallbacks++;
      std::vector<float> hh(w.n), uh(w.n), vh(w.n), bh(w.n);
      size_t fb = w.n * sizeof(float);
      HIPCHK(hipMemcpyAsync(hh.data(), w.planes[w.cur][0], fb, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(uh.data(), w.planes[w.cur][1], fb, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(vh.data(), w.planes[w.cur][2], fb, hipMemcpyDeviceToHost, st));
      HIPCHK(hipMemcpyAsync(bh.data(), w.bed, fb, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      plans[k] = plan_step(hh.data(), uh.data(), vh.data(), bh.data(), w.res, w.cell, dt, 0, 0.0f, 0.0f, kEps,
                           8.0f, 0);
      continue;
    }
    int rc = plan_nsub_begin(w.planes[w.cur][1], w.planes[w.cur][2], w.res, w.cell, dt, b->h_stats[k], b->ps[k],
                             &plans[k], st);
    if (rc < 0) return rc;
    wait = wait || b->ps[k].maxs2_pending;
  }
  if (wait) HIPCHK(hipStreamSynchronize(st));
  int nsub_max = 0;
  for (size_t k = 0; k < n; k++) {
    if (!host[k]) plan_nsub_end(b->ps[k], b->w[k].cell, dt, &plans[k]);
    if (!plans[k].boxes.empty()) nsub_max = std::max(nsub_max, plans[k].nsub);
  }
  if (nsub_max == 0) {
    b->last_ms = 0.0f;
    return 0;
  }
  // Launch s holds substep s of every window that has one.  A plan the fused
  // kernel does not cover (touching boxes, more than eight, or a box the
  // reference sweeps in place: under 48 x 32 cells) runs on the stage
  // kernels after the launches.
  const bool exact = b->path == SILT_PATH_FUSED_EXACT;
  std::vector<size_t> staged;
  for (size_t k = 0; k < n; k++) {
    std::vector<Box>& bx = plans[k].boxes;
    bool fits = bx.size() <= 8;
    for (size_t i = 0; i < bx.size() && fits; i++)
      for (size_t j = i + 1; j < bx.size() && fits; j++)
        fits = bx[i].i1 + 1 < bx[j].i0 || bx[j].i1 + 1 < bx[i].i0 || bx[i].j1 + 1 < bx[j].j0 || bx[j].j1 + 1 < bx[i].j0;
    for (size_t i = 0; i < bx.size() && fits; i++)
      fits = b->w[k].vort <= 0.0f || reference_uses_jacobi(bx[i], b->w[k].cell, 0, true);
    if (!bx.empty() && !fits) staged.push_back(k);
  }
  auto is_staged = [&](size_t k) { return std::find(staged.begin(), staged.end(), k) != staged.end(); };
  std::vector<FusedLaunch> launches((size_t)nsub_max);
  std::vector<int> cur(n);
  for (size_t k = 0; k < n; k++) cur[k] = b->w[k].cur;
  for (int s = 0; s < nsub_max; s++) {
    std::vector<FusedLayer> wins;
    for (size_t k = 0; k < n; k++) {
      const StepPlan& p = plans[k];
      if (p.boxes.empty() || p.nsub <= s || is_staged(k)) continue;
      Window& w = b->w[k];
      FusedLayer f{};
      f.in.h = w.planes[cur[k]][0];
      f.in.u = w.planes[cur[k]][1];
      f.in.v = w.planes[cur[k]][2];
      f.in.bed = w.bed;
      f.in.nu = nullptr;
      f.in.gate = nullptr;
      f.res = w.res;
      f.ho = w.planes[1 - cur[k]][0];
      f.uo = w.planes[1 - cur[k]][1];
      f.vo = w.planes[1 - cur[k]][2];
      silt_drive d{SILT_DRIVE_GRAVITY, 0.0f, 0.0f, w.vort};
      f.a = make_args(w.res, w.cell, dt / (float)p.nsub, d, kEps, 0, true, true, w.nu);
      f.boxes = p.boxes;
      f.stats = s == p.nsub - 1 ? b->stats + k : nullptr;
      f.totals = b->totals + k;
      wins.push_back(f);
      cur[k] = 1 - cur[k];
    }
    if (wins.empty()) {
      launches.resize((size_t)s);
      break;
    }
    int rc = fused_build(wins.data(), (int)wins.size(), exact, b->rows, &launches[(size_t)s],
                         b->xcd_on ? b->xcd_w : nullptr);
    if (rc != 0) return rc;
    const FusedLaunch& l = launches[(size_t)s];
    Slot& sl = b->slot[s];
    if ((rc = grow(sl, l.descs.size(), l.tile_window.size(), l.rows.size())) != 0) return rc;
    memcpy(sl.h_table, l.descs.data(), l.descs.size());
    if (!l.tile_window.empty()) memcpy(sl.h_tiles, l.tile_window.data(), l.tile_window.size() * sizeof(int));
    if (!l.rows.empty()) memcpy(sl.h_rows, l.rows.data(), l.rows.size() * sizeof(int));
  }
  // The launch shape: a new graph when it changes.
  std::vector<long long> key;
  for (const FusedLaunch& l : launches) {
    key.push_back(l.tiles);
    key.push_back(l.nwin);
    key.push_back(((long long)l.exact << 3) | ((long long)l.full << 2) | ((long long)l.project << 1) | l.wind);
    key.push_back((long long)l.rows.size());
    const Slot& sl = b->slot[&l - launches.data()];
    for (const void* ptr : {(const void*)sl.d_table, (const void*)sl.d_tiles, (const void*)sl.h_table,
                            (const void*)sl.h_tiles, (const void*)sl.d_rows, (const void*)sl.h_rows})
      key.push_back((long long)(size_t)ptr);
  }
  HIPCHK(hipEventRecord(b->ev0, st));
  if (launches.empty()) {
    stats_reset(b->stats, (int)n, st);
  } else if (b->use_graph) {
    if (!b->exec || key != b->graph_key) {
      drop_graph(b);
      HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc = enqueue_step(b, launches);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(st, &g);
      if (rc != 0) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess || !g) return SILT_EDEVICE;
      b->graph = g;
      if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        drop_graph(b);
        return SILT_EDEVICE;
      }
      b->graph_key = key;
      b->graphs++;
    }
    HIPCHK(hipGraphLaunch(b->exec, st));
  } else {
    int rc = enqueue_step(b, launches);
    if (rc != 0) return rc;
  }
  for (size_t k : staged) {
    Window& w = b->w[k];
    Workspace* ws = workspace_for(w.res);
    if (!ws) return SILT_ENOMEM;
    SheetPlanes sp = ws->planes();
    sp.h = w.planes[w.cur][0];
    sp.u = w.planes[w.cur][1];
    sp.v = w.planes[w.cur][2];
    sp.bed = w.bed;
    silt_sheet_opts o;
    silt_sheet_opts_default(&o);
    o.path = SILT_PATH_EXACT;
    silt_drive d{SILT_DRIVE_GRAVITY, 0.0f, 0.0f, w.vort};
    int rc = run_plan(sp, plans[k], w.res, w.cell, dt, d, o, true, true, w.nu, st, b->totals + k);
    if (rc < 0) return rc;
    stats_compute(sp.h, sp.u, sp.v, w.res, kEps, b->stats + k, st);
  }
  HIPCHK(hipEventRecord(b->ev1, st));
  HIPCHK(hipEventSynchronize(b->ev1));
  HIPCHK(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
  for (size_t k = 0; k < n; k++) b->w[k].cur = cur[k];
  return nsub_max;
}

int silt_batch_ledger(silt_batch* b, size_t k, silt_layer_ledger_t* out) {
  if (!b || k >= b->w.size() || !out) return SILT_EINVAL;
  SheetStats s;
  SheetTotals t;
  HIPCHK(hipMemcpyAsync(&s, b->stats + k, sizeof s, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipMemcpyAsync(&t, b->totals + k, sizeof t, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  const Window& w = b->w[k];
  double area = (double)w.cell * (double)w.cell;
  out->volume_m3 = s.vol * area;
  out->initial_m3 = w.vol0 * area;
  out->cut_m3 = t.cut * area;
  out->residue_m3 = t.residue * area;
  out->gap_m3 = (s.vol + t.cut - t.residue - w.vol0) * area;
  return SILT_OK;
}

float silt_batch_last_kernel_ms(const silt_batch* b) { return b ? b->last_ms : 0.0f; }
int silt_batch_last_fallbacks(const silt_batch* b) { return b ? b->last_fallbacks : SILT_EINVAL; }
int silt_batch_graphs_captured(const silt_batch* b) { return b ? b->graphs : SILT_EINVAL; }

int silt_batch_use_graph(silt_batch* b, int on) {
  if (!b) return SILT_EINVAL;
  b->use_graph = on != 0;
  if (!b->use_graph) drop_graph(b);
  return SILT_OK;
}

int silt_batch_set_rows(silt_batch* b, int rows) {
  if (!b || rows < 0) return SILT_EINVAL;
  b->rows = rows;
  return SILT_OK;
}

int silt_batch_autotune(silt_batch* b, float dt, const int* rows, size_t n, int reps, int* best_rows,
                        float* best_ms) {
  if (!b || !(dt > 0.0f) || reps < 1 || reps > 1000 || (n && !rows)) return SILT_EINVAL;
  std::vector<int> cand;
  if (rows && n) {
    for (size_t k = 0; k < n; k++)
      if (rows[k] >= 0) cand.push_back(rows[k]);
  } else {
    cand = {0, 16, 24, 32, 48, 64, 96, 128, 192, 256};
  }
  if (cand.empty()) return SILT_EINVAL;
  const size_t nw = b->w.size();
  hipStream_t st = b->stream;
  // Saved state: every window's current planes, statistics and ledger lines.
  size_t total = 0;
  for (const Window& w : b->w) total += 3 * w.n;
  float* save = nullptr;
  if (hipMalloc(&save, total * sizeof(float)) != hipSuccess) return SILT_ENOMEM;
  std::vector<int> cur0(nw);
  std::vector<SheetStats> stats0(nw);
  std::vector<SheetTotals> totals0(nw);
  int rc = 0;
  auto copy_planes = [&](bool to_save) -> int {
    size_t off = 0;
    for (size_t k = 0; k < nw; k++) {
      const Window& w = b->w[k];
      size_t fb = w.n * sizeof(float);
      for (int c = 0; c < 3; c++) {
        float* dev = w.planes[cur0[k]][c];
        if (hipMemcpyAsync(to_save ? save + off : dev, to_save ? dev : save + off, fb, hipMemcpyDeviceToDevice,
                           st) != hipSuccess)
          return SILT_EDEVICE;
        off += w.n;
      }
    }
    return 0;
  };
  for (size_t k = 0; k < nw; k++) cur0[k] = b->w[k].cur;
  rc = copy_planes(true);
  if (rc == 0 && (hipMemcpyAsync(stats0.data(), b->stats, nw * sizeof(SheetStats), hipMemcpyDeviceToHost, st) !=
                      hipSuccess ||
                  hipMemcpyAsync(totals0.data(), b->totals, nw * sizeof(SheetTotals), hipMemcpyDeviceToHost, st) !=
                      hipSuccess ||
                  hipStreamSynchronize(st) != hipSuccess))
    rc = SILT_EDEVICE;
  auto restore = [&]() -> int {
    for (size_t k = 0; k < nw; k++) b->w[k].cur = cur0[k];
    int r = copy_planes(false);
    if (r) return r;
    HIPCHK(hipMemcpyAsync(b->stats, stats0.data(), nw * sizeof(SheetStats), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b->totals, totals0.data(), nw * sizeof(SheetTotals), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
  };
  const int rows0 = b->rows;
  float best = INFINITY;
  int best_c = rows0;
  for (size_t c = 0; c < cand.size() && rc == 0; c++) {
    b->rows = cand[c];
    rc = silt_batch_step(b, dt);
    double sum = 0.0;
    for (int k = 0; k < reps && rc >= 0; k++) {
      rc = silt_batch_step(b, dt);
      sum += b->last_ms;
    }
    int rr = restore();
    if (rc >= 0) rc = rr;
    if (rc < 0) break;
    rc = 0;
    float mean = (float)(sum / reps);
    if (mean < best) {
      best = mean;
      best_c = cand[c];
    }
  }
  b->rows = rc == 0 ? best_c : rows0;
  // Second knob: with XCD weights given, the best height with and without
  // them; the weights stay on only when faster.
  if (rc == 0 && b->xcd_have) {
    const bool on0 = b->xcd_on;
    b->xcd_on = !on0;
    rc = silt_batch_step(b, dt);
    double sum = 0.0;
    for (int k = 0; k < reps && rc >= 0; k++) {
      rc = silt_batch_step(b, dt);
      sum += b->last_ms;
    }
    int rr = restore();
    if (rc >= 0) rc = rr;
    if (rc >= 0) {
      rc = 0;
      const float mean = (float)(sum / reps);
      if (mean < best)
        best = mean;
      else
        b->xcd_on = on0;
    } else {
      b->xcd_on = on0;
    }
  }
  (void)hipFree(save);
  if (rc < 0) return rc;
  if (best_rows) *best_rows = best_c;
  if (best_ms) *best_ms = best;
  return SILT_OK;
}

int silt_batch_set_xcd_weights(silt_batch* b, const float* w) {
  if (!b) return SILT_EINVAL;
  if (!w) {
    b->xcd_have = b->xcd_on = false;
    for (int x = 0; x < kXcds; x++) b->xcd_w[x] = 1.0f;
    return SILT_OK;
  }
  for (int x = 0; x < kXcds; x++)
    if (!(w[x] >= 0.25f && w[x] <= 4.0f)) return SILT_EINVAL;
  for (int x = 0; x < kXcds; x++) b->xcd_w[x] = w[x];
  b->xcd_have = b->xcd_on = true;
  return SILT_OK;
}

int silt_batch_xcd_weights_on(const silt_batch* b) { return b ? (b->xcd_on ? 1 : 0) : SILT_EINVAL; }

}  // extern "C"

int silt_plan_boxes_debug(const float* const* h, const int* res, size_t n, float cell, float eps, int* counts,
                          int* boxes, size_t cap) {
  if (!h || !res || !counts || !boxes || n == 0) return SILT_EINVAL;
  for (size_t k = 0; k < n; k++)
    if (!h[k] || res[k] <= 0) return SILT_EINVAL;
  hipStream_t st = nullptr;
  HIPCHK(hipStreamCreate(&st));
  std::vector<float*> d(n, nullptr);
  SheetStats* ds = nullptr;
  std::vector<SheetStats> hs(n);
  PlanLabeller lab;
  int rc = SILT_OK;
  auto done = [&](int r) {
    (void)hipStreamSynchronize(st);
    for (float* q : d)
      if (q) (void)hipFree(q);
    if (ds) (void)hipFree(ds);
    plan_labeller_free(lab);
    (void)hipStreamDestroy(st);
    return r;
  };
  size_t most = 0;
  for (size_t k = 0; k < n; k++) most = std::max(most, (size_t)res[k] * res[k]);
  float* zero = nullptr;
  if (hipMalloc(&ds, n * sizeof(SheetStats)) != hipSuccess || hipMalloc(&zero, most * sizeof(float)) != hipSuccess)
    return done(SILT_ENOMEM);
  d.push_back(zero);
  if (hipMemsetAsync(zero, 0, most * sizeof(float), st) != hipSuccess) return done(SILT_EDEVICE);
  for (size_t k = 0; k < n; k++) {
    const size_t cells = (size_t)res[k] * res[k];
    if (hipMalloc(&d[k], cells * sizeof(float)) != hipSuccess) return done(SILT_ENOMEM);
    if (hipMemcpyAsync(d[k], h[k], cells * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess)
      return done(SILT_EDEVICE);
    stats_reset(ds + k, 1, st);
    stats_compute(d[k], zero, zero, res[k], eps, ds + k, st);
  }
  if (hipMemcpyAsync(hs.data(), ds, n * sizeof(SheetStats), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return done(SILT_EDEVICE);
  std::vector<PlanSheet> sheets(n);
  for (size_t k = 0; k < n; k++) sheets[k] = PlanSheet{d[k], res[k], cell, &hs[k]};
  std::vector<StepPlan> plans(n);
  std::vector<char> host(n, 0);
  rc = plan_boxes_many(lab, sheets.data(), (int)n, 1.0f / 120.0f, eps, plans.data(), host.data(), st);
  if (rc < 0) return done(rc);
  for (size_t k = 0; k < n; k++) {
    counts[k] = host[k] ? -1 : (int)plans[k].boxes.size();
    for (size_t b = 0; b < plans[k].boxes.size() && b < cap; b++) {
      const Box& x = plans[k].boxes[b];
      int* o = boxes + 4 * (cap * k + b);
      o[0] = x.i0;
      o[1] = x.i1;
      o[2] = x.j0;
      o[3] = x.j1;
    }
  }
  return done(SILT_OK);
}
