// C ABI of the stateless sheet step (include/silt_sheet.h): step_sheet on host
// slices and on device pointers.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <map>
#include <mutex>
#include <vector>

#include "plan.hpp"
#include "sheet_device.hpp"
#include "sheet_fused.hpp"
#include "silt_sheet.h"
#include "workspace.hpp"

#define HIPCHK(x)                                   \
  do {                                              \
    if ((x) != hipSuccess) return SILT_EDEVICE;     \
  } while (0)

namespace silt {

StepArgs make_args(int res, float cell, float dt, const silt_drive& d, float eps, int banks, bool gate_uniform,
                   bool nu_uniform, float nu0) {
  StepArgs a{};
  a.res = res;
  a.cell = cell;
  a.dt = dt;
  a.wind = d.kind == SILT_DRIVE_WIND;
  a.wx = a.wind ? d.x : 0.0f;
  a.wz = a.wind ? d.z : 0.0f;
  a.vort = d.vorticity;
  a.eps = eps;
  a.project = !a.wind && cell > kFineCell + 1.0e-6f;
  a.lim = a.wind ? fmaxf(hypotf(a.wx, a.wz), kFilmSpeedCap) : kFilmSpeedCap;
  a.blend = 1.0f - expf(-dt * 1.6f);
  a.banks = banks;
  a.gate_uniform = gate_uniform;
  a.nu_uniform = nu_uniform;
  a.nu0 = nu0;
  a.mob0 = sheet_mobility(nu0);
  a.ov0 = overload_of(nu0);
  a.decay0 = expf(-sheet_damp(nu0) * dt);
  return a;
}

// Runs a planned step on device planes.  Exact path: stage kernels per box,
// boxes in order, as the reference loops them.  Fast path: the fused kernel.
int run_plan(const SheetPlanes& s, const StepPlan& plan, int res, float cell, float dt, const silt_drive& d,
             const silt_sheet_opts& o, bool gate_uniform, bool nu_uniform, float nu0, hipStream_t st,
             SheetTotals* book) {
  if (plan.boxes.empty() || plan.nsub <= 0) return 0;
  float dts = dt / (float)plan.nsub;
  bool wind = d.kind == SILT_DRIVE_WIND;
  if (s.solid) {
    // Walls: the stage kernels, whatever the path (the streaming kernel has
    // no walled form).  The parallel-kick form where the reference takes it.
    StepArgs a = make_args(res, cell, dts, d, o.eps, 0, gate_uniform, nu_uniform, nu0);
    a.book = book;
    for (int k = 0; k < plan.nsub; k++)
      for (const Box& b : plan.boxes) {
        bool par = !wind && o.row_split != 0 && b.j1 - b.j0 >= 47 && b.i1 - b.i0 >= 31;
        exact_substep_walled(s, b, a, par, st);
      }
    return plan.nsub;
  }
  auto gs_for = [&](const Box& b) {
    if (o.confine == SILT_CONFINE_JACOBI || d.vorticity <= 0.0f) return false;
    return !reference_uses_jacobi(b, cell, wind, o.row_split != 0);
  };
  if (o.path == SILT_PATH_FAST || o.path == SILT_PATH_FUSED_EXACT) {
    // The streaming kernel sweeps confinement from a snapshot; a box the
    // reference sweeps in place, a pace-3 step or overlapping boxes take the
    // stage kernels instead.
    bool any_gs = false;
    for (const Box& b : plan.boxes) any_gs = any_gs || gs_for(b);
    if (!any_gs) {
      int rc = fused_step(s, plan, res, cell, dt, d, o, gate_uniform, nu_uniform, nu0, st, book);
      if (rc != SILT_EUNSUPPORTED) return rc;
    }
  }
  if (plan.pace3) {
    StepArgs full = make_args(res, cell, dt, d, o.eps, o.banks, gate_uniform, nu_uniform, nu0);
    StepArgs slice = make_args(res, cell, dts, d, o.eps, o.banks, gate_uniform, nu_uniform, nu0);
    slice.banks = 0;
    full.book = slice.book = book;
    for (const Box& b : plan.boxes) {
      exact_substep(s, b, full, false, gs_for(b), st);
      for (int k = 0; k < plan.nsub; k++) exact_transport(s, b, slice, dts, st);
      exact_dry_cut(s, b, slice, st);
    }
    return plan.nsub;
  }
  StepArgs a = make_args(res, cell, dts, d, o.eps, o.banks, gate_uniform, nu_uniform, nu0);
  a.book = book;
  for (int k = 0; k < plan.nsub; k++)
    for (const Box& b : plan.boxes) exact_substep(s, b, a, true, gs_for(b), st);
  return plan.nsub;
}

// A step of one layer whose planes are already on the device: the plan reads
// depth, velocity and bed back to the host (blocking), then the kernels are
// enqueued on st.  The workspace scratch planes in s must be valid.
int step_device_planes(const SheetPlanes& s, int res, float cell, float dt, const silt_drive& d,
                       const silt_sheet_opts& o, bool gate_uniform, bool nu_uniform, float nu0, hipStream_t st,
                       SheetTotals* book) {
  size_t n = (size_t)res * res;
  std::vector<float> hh(n), uh(n), vh(n), bh(n);
  size_t fb = n * sizeof(float);
  HIPCHK(hipMemcpyAsync(hh.data(), s.h, fb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(uh.data(), s.u, fb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(vh.data(), s.v, fb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(bh.data(), s.bed, fb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  StepPlan plan = plan_step(hh.data(), uh.data(), vh.data(), bh.data(), res, cell, dt, d.kind == SILT_DRIVE_WIND, d.x,
                            d.z, o.eps, o.sub_cap, o.pace, s.solid != nullptr);
  if (plan.boxes.empty()) return 0;
  return run_plan(s, plan, res, cell, dt, d, o, gate_uniform, nu_uniform, nu0, st, book);
}

bool valid_opts(const silt_sheet_opts& o) {
  return (o.pace >= 0 && o.pace <= 255) && (o.confine == SILT_CONFINE_REFERENCE || o.confine == SILT_CONFINE_JACOBI) &&
         (o.path == SILT_PATH_EXACT || o.path == SILT_PATH_FAST || o.path == SILT_PATH_FUSED_EXACT);
}

}  // namespace silt

using namespace silt;

extern "C" {

void silt_sheet_opts_default(silt_sheet_opts* o) {
  if (!o) return;
  o->eps = 1.0e-6f;
  o->banks = 0;
  o->sub_cap = 8.0f;
  o->pace = 0;
  o->row_split = 1;
  o->confine = SILT_CONFINE_REFERENCE;
  o->path = SILT_PATH_EXACT;
}

const char* silt_version(void) { return "silt_amd 0.1 gfx950"; }

int silt_set_chunk_rows(size_t res, int rows) {
  if (res < 1 || res > 65536 || rows < 0) return SILT_EINVAL;
  fused_set_chunk_rows_for((int)res, rows);
  return SILT_OK;
}

int silt_chunk_rows(size_t res) {
  if (res < 1 || res > 65536) return SILT_EINVAL;
  return fused_pinned_chunk_rows((int)res);
}

int silt_step_sheet_ex(float* h, float* u, float* v, const float* bed, const float* nu, const uint8_t* gate,
                       size_t len, size_t res, float cell, float dt, silt_drive drive, const silt_sheet_opts* opts) {
  if (!h || !u || !v || !bed || !nu || !gate) return SILT_EINVAL;
  if (drive.kind != SILT_DRIVE_GRAVITY && drive.kind != SILT_DRIVE_WIND) return SILT_EINVAL;
  silt_sheet_opts o;
  if (opts)
    o = *opts;
  else
    silt_sheet_opts_default(&o);
  if (!valid_opts(o)) return SILT_EINVAL;
  // The reference's silent early returns (src/fluid_core.rs:873-880).
  if (!(dt > 0.0f) || res == 0 || len < res * res) return 0;
  size_t n = res * res;
  StepPlan plan = plan_step(h, u, v, bed, (int)res, cell, dt, drive.kind == SILT_DRIVE_WIND, drive.x, drive.z, o.eps,
                            o.sub_cap, o.pace);
  if (plan.boxes.empty()) return 0;
  bool gate_uniform = true;
  for (size_t k = 0; k < n && gate_uniform; k++) gate_uniform = gate[k] == 1;
  bool nu_uniform = true;
  for (size_t k = 1; k < n && nu_uniform; k++) nu_uniform = memcmp(&nu[k], &nu[0], sizeof(float)) == 0;
  Workspace* ws = workspace_for((int)res);
  if (!ws) return SILT_ENOMEM;
  std::lock_guard<std::mutex> lock(ws->mu);
  SheetPlanes s = ws->planes();
  hipStream_t st = ws->stream;
  size_t fb = n * sizeof(float);
  HIPCHK(hipMemcpyAsync(s.h, h, fb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s.u, u, fb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s.v, v, fb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync((void*)s.bed, bed, fb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync((void*)s.nu, nu, fb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync((void*)s.gate, gate, n, hipMemcpyHostToDevice, st));
  int nsub = run_plan(s, plan, (int)res, cell, dt, drive, o, gate_uniform, nu_uniform, nu[0], st);
  HIPCHK(hipMemcpyAsync(h, s.h, fb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(u, s.u, fb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(v, s.v, fb, hipMemcpyDeviceToHost, st));
  if (hipStreamSynchronize(st) != hipSuccess) return SILT_EDEVICE;
  if (hipGetLastError() != hipSuccess) return SILT_EDEVICE;
  return nsub;
}

int silt_step_sheet_walled(float* h, float* u, float* v, const float* bed, const float* nu, const uint8_t* gate,
                           const uint8_t* solid, size_t len, size_t res, float cell, float dt, silt_drive drive,
                           const silt_sheet_opts* opts) {
  if (!h || !u || !v || !bed || !nu || !gate || !solid) return SILT_EINVAL;
  if (drive.kind != SILT_DRIVE_GRAVITY && drive.kind != SILT_DRIVE_WIND) return SILT_EINVAL;
  silt_sheet_opts o;
  if (opts)
    o = *opts;
  else
    silt_sheet_opts_default(&o);
  if (!valid_opts(o)) return SILT_EINVAL;
  if (!(dt > 0.0f) || res == 0 || len < res * res) return 0;
  size_t n = res * res;
  StepPlan plan = plan_step(h, u, v, bed, (int)res, cell, dt, drive.kind == SILT_DRIVE_WIND, drive.x, drive.z, o.eps,
                            o.sub_cap, o.pace, true);
  if (plan.boxes.empty()) return 0;
  bool gate_uniform = true;
  for (size_t k = 0; k < n && gate_uniform; k++) gate_uniform = gate[k] == 1;
  bool nu_uniform = true;
  for (size_t k = 1; k < n && nu_uniform; k++) nu_uniform = memcmp(&nu[k], &nu[0], sizeof(float)) == 0;
  Workspace* ws = workspace_for((int)res);
  if (!ws) return SILT_ENOMEM;
  std::lock_guard<std::mutex> lock(ws->mu);
  SheetPlanes s = ws->planes();
  s.solid = ws->solid_stage();
  hipStream_t st = ws->stream;
  size_t fb = n * sizeof(float);
  HIPCHK(hipMemcpyAsync(s.h, h, fb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s.u, u, fb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s.v, v, fb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync((void*)s.bed, bed, fb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync((void*)s.nu, nu, fb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync((void*)s.gate, gate, n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync((void*)s.solid, solid, n, hipMemcpyHostToDevice, st));
  int nsub = run_plan(s, plan, (int)res, cell, dt, drive, o, gate_uniform, nu_uniform, nu[0], st);
  HIPCHK(hipMemcpyAsync(h, s.h, fb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(u, s.u, fb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(v, s.v, fb, hipMemcpyDeviceToHost, st));
  if (hipStreamSynchronize(st) != hipSuccess) return SILT_EDEVICE;
  if (hipGetLastError() != hipSuccess) return SILT_EDEVICE;
  return nsub;
}

int silt_step_sheet_walled_device(float* h, float* u, float* v, const float* bed, const float* nu, float nu_uniform,
                                  const uint8_t* gate, const uint8_t* solid, size_t res, float cell, float dt,
                                  silt_drive drive, const silt_sheet_opts* opts, void* hip_stream) {
  bool uniform = isfinite(nu_uniform);
  if (!h || !u || !v || !bed || !solid || (!nu && !uniform)) return SILT_EINVAL;
  if (drive.kind != SILT_DRIVE_GRAVITY && drive.kind != SILT_DRIVE_WIND) return SILT_EINVAL;
  silt_sheet_opts o;
  if (opts)
    o = *opts;
  else
    silt_sheet_opts_default(&o);
  if (!valid_opts(o)) return SILT_EINVAL;
  if (!(dt > 0.0f) || res == 0) return 0;
  Workspace* ws = workspace_for((int)res);
  if (!ws) return SILT_ENOMEM;
  std::lock_guard<std::mutex> lock(ws->mu);
  SheetPlanes s = ws->planes();
  s.h = h;
  s.u = u;
  s.v = v;
  s.bed = bed;
  s.nu = uniform ? s.nu : nu;
  s.gate = gate ? gate : s.gate;
  s.solid = solid;
  int nsub = step_device_planes(s, (int)res, cell, dt, drive, o, gate == nullptr, uniform, uniform ? nu_uniform : 0.0f,
                                (hipStream_t)hip_stream);
  if (nsub < 0) return nsub;
  if (hipStreamSynchronize((hipStream_t)hip_stream) != hipSuccess) return SILT_EDEVICE;
  if (hipGetLastError() != hipSuccess) return SILT_EDEVICE;
  return nsub;
}

int silt_displace_occupied(float* h, float* u, float* v, const uint8_t* solid, size_t len, size_t res) {
  if (!h || !u || !v || !solid) return SILT_EINVAL;
  if (res == 0 || len < res * res) return 0;
  size_t n = res * res;
  Workspace* ws = workspace_for((int)res);
  if (!ws) return SILT_ENOMEM;
  std::lock_guard<std::mutex> lock(ws->mu);
  SheetPlanes s = ws->planes();
  hipStream_t st = ws->stream;
  size_t fb = n * sizeof(float);
  uint8_t* dsolid = ws->solid_stage();
  HIPCHK(hipMemcpyAsync(s.h, h, fb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s.u, u, fb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s.v, v, fb, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dsolid, solid, n, hipMemcpyHostToDevice, st));
  displace_occupied(s.h, s.u, s.v, dsolid, (int)res, s.tg, s.ts, nullptr, st);
  HIPCHK(hipMemcpyAsync(h, s.h, fb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(u, s.u, fb, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(v, s.v, fb, hipMemcpyDeviceToHost, st));
  if (hipStreamSynchronize(st) != hipSuccess) return SILT_EDEVICE;
  if (hipGetLastError() != hipSuccess) return SILT_EDEVICE;
  return SILT_OK;
}

int silt_step_sheet(float* h, float* u, float* v, const float* bed, const float* nu, const uint8_t* gate, size_t len,
                    size_t res, float cell, float dt, silt_drive drive) {
  return silt_step_sheet_ex(h, u, v, bed, nu, gate, len, res, cell, dt, drive, nullptr);
}

int silt_step_sheet_device(float* h, float* u, float* v, const float* bed, const float* nu, float nu_uniform,
                           const uint8_t* gate, size_t res, float cell, float dt, silt_drive drive,
                           const silt_sheet_opts* opts, void* hip_stream) {
  bool uniform = isfinite(nu_uniform);
  if (!h || !u || !v || !bed || (!nu && !uniform)) return SILT_EINVAL;
  if (drive.kind != SILT_DRIVE_GRAVITY && drive.kind != SILT_DRIVE_WIND) return SILT_EINVAL;
  silt_sheet_opts o;
  if (opts)
    o = *opts;
  else
    silt_sheet_opts_default(&o);
  if (!valid_opts(o)) return SILT_EINVAL;
  if (!(dt > 0.0f) || res == 0) return 0;
  Workspace* ws = workspace_for((int)res);
  if (!ws) return SILT_ENOMEM;
  std::lock_guard<std::mutex> lock(ws->mu);
  SheetPlanes s = ws->planes();
  s.h = h;
  s.u = u;
  s.v = v;
  s.bed = bed;
  s.nu = uniform ? s.nu : nu;
  s.gate = gate ? gate : s.gate;
  int nsub = step_device_planes(s, (int)res, cell, dt, drive, o, gate == nullptr, uniform, uniform ? nu_uniform : 0.0f,
                                (hipStream_t)hip_stream);
  if (nsub < 0) return nsub;
  if (hipStreamSynchronize((hipStream_t)hip_stream) != hipSuccess) return SILT_EDEVICE;
  if (hipGetLastError() != hipSuccess) return SILT_EDEVICE;
  return nsub;
}

}  // extern "C"
