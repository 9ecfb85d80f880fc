// silt_rut_stamp_device_noalloc (include/silt_rut.h): the rut stamp of
// silt_rut_stamp_device for callers that may not allocate memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The stamp kernel reads its prepared contacts from device memory.  Here that
// buffer is `scratch`, a file-scope device array of kNoAllocContacts records
// (88 bytes each, 88 KiB per GPU, held until the process ends), and the host
// side prepares the contacts into a static array of the same size.  The
// array's device address is taken once per device with hipGetSymbolAddress
// and used for every copy and launch; nothing is allocated per call.
#include <hip/hip_runtime.h>

#include <mutex>

#include "rut.hpp"
#include "silt_rut.h"

using namespace silt;

namespace {

constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) RutContact scratch[SILT_NOALLOC_MAX_CONTACTS];

namespace {

std::mutex g_mu;
RutContact* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device
RutContact g_host[SILT_NOALLOC_MAX_CONTACTS];

// The current device's address of `scratch`, looked up on its first call.
// Called with g_mu held.
int scratch_address(RutContact** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return SILT_EDEVICE;
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SILT_EDEVICE;
    g_scratch[dev] = static_cast<RutContact*>(p);
  }
  *out = g_scratch[dev];
  return SILT_OK;
}

}  // namespace

extern "C" {

int silt_rut_stamp_device_noalloc(float* bed, float* placed, const float* initial, size_t res, float cell,
                                  float origin_x, float origin_z, const silt_tire_contact* contacts, size_t n, float dt,
                                  void* hip_stream) {
  if (!bed || !placed || !initial || (!contacts && n) || res < 3 || !(cell > 0.0f)) return SILT_EINVAL;
  if (n > (size_t)SILT_NOALLOC_MAX_CONTACTS) return SILT_ETOOBIG;
  if (n == 0 || !(dt > 0.0f)) return 0;
  std::lock_guard<std::mutex> lock(g_mu);
  RutContact* d = nullptr;
  int rc = scratch_address(&d);
  if (rc) return rc;
  rc = prepare_contacts_into(contacts, n, (int)res, cell, origin_x, origin_z, dt, g_host);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)hip_stream;
  rc = rut_stamp_prepared(bed, placed, initial, (int)res, cell, origin_x, origin_z, g_host, (int)n, d, st);
  if (rc) return rc;
  // Every call shares the one array: finish this call's stamp before the
  // next call rewrites it (as silt_rut_stamp_device does with its buffer).
  return hipStreamSynchronize(st) == hipSuccess ? 0 : SILT_EDEVICE;
}

}  // extern "C"
