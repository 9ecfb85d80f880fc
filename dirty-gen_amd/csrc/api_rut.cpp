// C ABI of the Bekker rut coupling (include/silt_rut.h).
#include <hip/hip_runtime.h>

#include <mutex>

#include "rut.hpp"
#include "silt_rut.h"

using namespace silt;

extern "C" {

float silt_bekker_load(float sinkage, float radius, float width, float k_eq, float n) {
  return bekker_load(sinkage, radius, width, k_eq, n);
}

float silt_rut_depth(float normal, float push, float radius, float width, float k_eq, float n) {
  return rut_depth_from_force(normal, push, radius, width, k_eq, n);
}

int silt_rut_stamp_device(float* bed, float* placed, const float* initial, size_t res, float cell, float origin_x,
                          float origin_z, const silt_tire_contact* contacts, size_t n, float dt, void* hip_stream) {
  if (!bed || !placed || !initial || (!contacts && n) || res < 3 || !(cell > 0.0f)) return SILT_EINVAL;
  if (n == 0 || !(dt > 0.0f)) return 0;
  static std::mutex mu;
  static RutContact* scratch = nullptr;
  static size_t scratch_n = 0;
  std::lock_guard<std::mutex> lock(mu);
  hipStream_t st = (hipStream_t)hip_stream;
  if (scratch_n < n) {
    if (scratch) {
      if (hipStreamSynchronize(st) != hipSuccess) return SILT_EDEVICE;
      (void)hipFree(scratch);
    }
    scratch = nullptr;
    scratch_n = 0;
    if (hipMalloc(&scratch, n * sizeof(RutContact)) != hipSuccess) return SILT_ENOMEM;
    scratch_n = n;
  }
  int rc = rut_stamp(bed, placed, initial, (int)res, cell, origin_x, origin_z, contacts, n, dt, scratch, st);
  if (rc) return rc;
  // The contact scratch is reused by the next call: finish this one first.
  return hipStreamSynchronize(st) == hipSuccess ? 0 : SILT_EDEVICE;
}

}  // extern "C"
