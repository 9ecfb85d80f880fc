// The two-layer FluidCore on the GPU (include/silt_core.h).  The host keeps
// the small decisions of a step (where the 2 cm patch goes, which coarse
// columns it holds, which tires shove which grid); every plane operation is
// a kernel on the handle's stream and the planes never leave the device,
// except the read back the sheet step's work-box plan takes.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "core.hpp"
#include "plan.hpp"
#include "rut.hpp"
#include "sheet_device.hpp"
#include "sheet_fused.hpp"
#include "silt_core.h"
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
  // Ground under the sheet (SoilField bed / placed / initial height on the
  // window grid), set by silt_core_set_ground and cut by the rut stamp.
  bool ground = false;
  float *soil_bed = nullptr, *soil_placed = nullptr, *soil_init = nullptr, *soil_rut = nullptr;
  RutContact* d_contacts = nullptr;
  hipStream_t st = nullptr;
  std::vector<void*> allocs;
};

namespace {

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
  core_restrict(c->c, c->f, c->d_hold, sp, ax, c->d_acc + 1, c->st);
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
    core_near_apply(c->f, c->near[c->near_cur], c->d_cols, (int)cols.size(), c->st);
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

int step_pair(silt_core* c, float dt, bool fine) {
  if (fine && !c->fine_on) return 0;
  int res = fine ? kFineRes : kWindowRes;
  float cell = fine ? kFineCell : kCoarseCell;
  Workspace* ws = workspace_for(res);
  if (!ws) return SILT_ENOMEM;
  silt_sheet_opts o;
  silt_sheet_opts_default(&o);
  o.path = c->path;
  o.row_split = 0;
  o.pace = fine ? c->fine_pace : 0;
  o.sub_cap = fine ? c->fine_sub_cap : 8.0f;
  o.confine = fine ? SILT_CONFINE_REFERENCE : c->coarse_confine;
  SheetPlanes s = ws->planes();
  const float nu_w = dynamic_viscosity(3, 1.0f);
  // Water layer.
  if (!fine) core_soften(c->c.uw, c->c.vw, res, c->st);
  s.h = fine ? c->f.water : c->c.water;
  s.u = fine ? c->f.uw : c->c.uw;
  s.v = fine ? c->f.vw : c->c.vw;
  s.bed = fine ? c->f.bed : c->c.bed;
  silt_drive dw{SILT_DRIVE_GRAVITY, 0.0f, 0.0f, 0.08f};
  int rc = step_device_planes(s, res, cell, dt, dw, o, true, true, nu_w, c->st);
  if (rc < 0) return rc;
  // Slurry layer at its own viscosity.
  float* nu = fine ? c->nu_f : c->nu_c;
  core_nu(nu, fine ? c->f.grain : c->c.grain, fine ? c->f.moisture : c->c.moisture, res * res, c->st);
  if (!fine) core_soften(c->c.us, c->c.vs, res, c->st);
  s.h = fine ? c->f.slurry : c->c.slurry;
  s.u = fine ? c->f.us : c->c.us;
  s.v = fine ? c->f.vs : c->c.vs;
  s.bed = fine ? c->f.bed_s : c->c.bed;
  s.nu = nu;
  silt_drive ds{SILT_DRIVE_GRAVITY, 0.0f, 0.0f, 0.05f};
  rc = step_device_planes(s, res, cell, dt, ds, o, true, false, 0.0f, c->st);
  return rc < 0 ? rc : 0;
}

void peel(silt_core* c, const PushList& pl, bool fine, bool skip_held) {
  GridGeo g{fine ? kFineRes : kWindowRes, fine ? kFineCell : kCoarseCell, fine ? c->fine_origin[0] : c->origin[0],
            fine ? c->fine_origin[1] : c->origin[1]};
  GridGeo cg{kWindowRes, kCoarseCell, c->origin[0], c->origin[1]};
  core_peel(g, fine ? c->f.water : c->c.water, fine ? c->f.uw : c->c.uw, fine ? c->f.vw : c->c.vw, pl, 0, 0, g.res,
            g.res, c->d_hold, cg, skip_held, c->d_ev, c->d_nev, kSprayCap, c->d_acc, c->st);
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

}  // namespace

extern "C" {

silt_core* silt_core_create(float center_x, float center_z, int path) {
  if (path != SILT_PATH_EXACT && path != SILT_PATH_FAST && path != SILT_PATH_FUSED_EXACT) return nullptr;
  silt_core* c = new silt_core();
  c->path = path;
  float extent = (float)kWindowRes * kCoarseCell;
  c->origin[0] = floorf((center_x - extent * 0.5f) / kCoarseCell) * kCoarseCell;
  c->origin[1] = floorf((center_z - extent * 0.5f) / kCoarseCell) * kCoarseCell;
  c->hold.assign(WN, 0);
  bool ok = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) == hipSuccess;
  float** cp[8] = {&c->c.water, &c->c.slurry, &c->c.uw, &c->c.vw, &c->c.us, &c->c.vs, &c->c.bed, &c->c.moisture};
  for (auto p : cp) ok = ok && dalloc(c, p, WN);
  float** fp[9] = {&c->f.water, &c->f.slurry, &c->f.uw, &c->f.vw, &c->f.us, &c->f.vs, &c->f.bed, &c->f.bed_s,
                   &c->f.moisture};
  for (auto p : fp) ok = ok && dalloc(c, p, FN);
  ok = ok && dalloc(c, &c->c.grain, WN) && dalloc(c, &c->f.grain, FN) && dalloc(c, &c->d_hold, WN) &&
       dalloc(c, &c->nu_c, WN) && dalloc(c, &c->nu_f, FN) && near_alloc(c, &c->near[0]) && near_alloc(c, &c->near[1]) &&
       dalloc(c, &c->d_cols, 1024) && dalloc(c, &c->d_ev, kSprayCap) && dalloc(c, &c->d_nev, 1) &&
       dalloc(c, &c->d_acc, 4) && dalloc(c, &c->soil_bed, WN) && dalloc(c, &c->soil_placed, WN) &&
       dalloc(c, &c->soil_init, WN) && dalloc(c, &c->soil_rut, WN) && dalloc(c, &c->d_contacts, kMaxPush);
  if (ok) {
    // Grain::Other everywhere (FluidCore::new, src/fluid_core.rs:4979-4980).
    ok = hipMemset(c->c.grain, 4, WN) == hipSuccess && hipMemset(c->f.grain, 4, FN) == hipSuccess;
    core_near_clear(c->near[0], c->st);
    core_near_clear(c->near[1], c->st);
    ok = ok && hipStreamSynchronize(c->st) == hipSuccess;
  }
  if (!ok) {
    silt_core_destroy(c);
    return nullptr;
  }
  return c;
}

void silt_core_destroy(silt_core* c) {
  if (!c) return;
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (void* p : c->allocs) (void)hipFree(p);
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
  if (restrict_fine(c)) return SILT_EDEVICE;
  std::vector<float> s(WN), m(WN);
  std::vector<uint8_t> g(WN);
  HIPCHK(hipMemcpyAsync(s.data(), c->c.slurry, WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(g.data(), c->c.grain, WN, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int k = 0; k < WN; k++) {
    m[k] = clampf(moisture[k], 0.0f, 1.0f);
    if (s[k] > 1.0e-5f) g[k] = grain[k];
  }
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
  std::vector<CorePush> pushes;
  pushes.swap(c->pushes);
  PushList pl = push_list(pushes);
  hipStream_t st = c->st;
  HIPCHK(hipMemsetAsync(c->d_nev, 0, sizeof(unsigned), st));
  HIPCHK(hipMemsetAsync(c->d_acc, 0, sizeof(double), st));
  int rc = refine(c, pushes, dt);
  if (rc) return rc;
  std::vector<CorePush> fine_p, coarse_p;
  for (const CorePush& p : pushes)
    if (cell_is_refined(c, p.x, p.z)) fine_p.push_back(p);
  shove(c, fine_p, true);
  if ((rc = step_pair(c, dt, true))) return rc;
  if (c->fine_on) peel(c, pl, true, false);
  if ((rc = restrict_fine(c))) return rc;
  for (const CorePush& p : pushes)
    if (!fine_held_at(c, p.x, p.z)) coarse_p.push_back(p);
  shove(c, coarse_p, false);
  if ((rc = step_pair(c, dt, false))) return rc;
  peel(c, pl, false, true);
  unsigned nev = 0;
  double vol = 0.0;
  HIPCHK(hipMemcpyAsync(&nev, c->d_nev, sizeof nev, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&vol, c->d_acc, sizeof vol, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
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
  Workspace* ws = workspace_for(kWindowRes);
  if (!ws) return SILT_ENOMEM;
  SheetStats* s = nullptr;
  if (hipMalloc(&s, 4 * sizeof(SheetStats)) != hipSuccess) return SILT_ENOMEM;
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
  HIPCHK(hipStreamSynchronize(c->st));
  (void)hipFree(s);
  double ca = (double)kCoarseCell * kCoarseCell, fa = (double)kFineCell * kFineCell;
  out->water_m3 = h[0].vol * ca + (c->fine_on ? h[2].vol * fa : 0.0);
  out->slurry_m3 = h[1].vol * ca + (c->fine_on ? h[3].vol * fa : 0.0);
  out->frozen_water_m3 = acc[1];
  out->frozen_slurry_m3 = acc[2];
  out->spray_m3 = c->spray_m3;
  out->total_m3 = out->water_m3 + out->slurry_m3 + out->frozen_water_m3 + out->frozen_slurry_m3 + out->spray_m3;
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
  if (c->ground && !soil_bed && !soil_placed && !soil_rut) {
    // The core's own ground: publish it without a round trip through the host.
    float* d_out = nullptr;
    if (hipMalloc(&d_out, (size_t)WN * 10 * sizeof(float)) != hipSuccess) return SILT_ENOMEM;
    core_publish(c->c, c->soil_bed, c->soil_placed, c->soil_rut, d_out, d_out + 4 * WN, d_out + 6 * WN, c->st);
    if (ground_rut4) HIPCHK(hipMemcpyAsync(ground_rut4, d_out, 4 * WN * 4, hipMemcpyDeviceToHost, c->st));
    if (film2) HIPCHK(hipMemcpyAsync(film2, d_out + 4 * WN, 2 * WN * 4, hipMemcpyDeviceToHost, c->st));
    if (flow4) HIPCHK(hipMemcpyAsync(flow4, d_out + 6 * WN, 4 * WN * 4, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    (void)hipFree(d_out);
    return SILT_OK;
  }
  float *d_out = nullptr, *d_soil = nullptr;
  if (hipMalloc(&d_out, (size_t)WN * 10 * sizeof(float)) != hipSuccess) return SILT_ENOMEM;
  if (hipMalloc(&d_soil, (size_t)WN * 3 * sizeof(float)) != hipSuccess) {
    (void)hipFree(d_out);
    return SILT_ENOMEM;
  }
  const float* soils[3] = {soil_bed, soil_placed, soil_rut};
  const float* dsoil[3] = {nullptr, nullptr, nullptr};
  for (int k = 0; k < 3; k++)
    if (soils[k]) {
      HIPCHK(hipMemcpyAsync(d_soil + (size_t)k * WN, soils[k], WN * 4, hipMemcpyHostToDevice, c->st));
      dsoil[k] = d_soil + (size_t)k * WN;
    }
  core_publish(c->c, dsoil[0], dsoil[1], dsoil[2], d_out, d_out + 4 * WN, d_out + 6 * WN, c->st);
  if (ground_rut4) HIPCHK(hipMemcpyAsync(ground_rut4, d_out, 4 * WN * 4, hipMemcpyDeviceToHost, c->st));
  if (film2) HIPCHK(hipMemcpyAsync(film2, d_out + 4 * WN, 2 * WN * 4, hipMemcpyDeviceToHost, c->st));
  if (flow4) HIPCHK(hipMemcpyAsync(flow4, d_out + 6 * WN, 4 * WN * 4, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  (void)hipFree(d_out);
  (void)hipFree(d_soil);
  return SILT_OK;
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

}  // extern "C"
