// Device workspace of the stateless entry points: staging planes for host
// slices and the scratch planes of one substep, cached per (device, res).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>

#include "sheet_device.hpp"

namespace silt {

struct Workspace {
  int device = 0;
  int res = 0;
  size_t n = 0;
  float* fbuf = nullptr;    // 5 staging planes + 8 scratch planes
  uint8_t* gate = nullptr;  // staging gate plane, then the staging wall mask
  hipStream_t stream = nullptr;
  std::mutex mu;

  uint8_t* solid_stage() const { return gate + n; }

  SheetPlanes planes() const {
    SheetPlanes s{};
    s.h = fbuf;
    s.u = fbuf + n;
    s.v = fbuf + 2 * n;
    s.bed = fbuf + 3 * n;
    s.nu = fbuf + 4 * n;
    s.gate = gate;
    s.eta = fbuf + 5 * n;
    s.tg = fbuf + 6 * n;
    s.ts = fbuf + 7 * n;
    s.p = fbuf + 8 * n;
    s.su = fbuf + 9 * n;
    s.sv = fbuf + 10 * n;
    s.bsrc = fbuf + 11 * n;
    s.h1 = fbuf + 12 * n;
    return s;
  }
};

constexpr int kWorkspacePlanes = 13;

// The calling thread's current device's workspace for grids of res x res,
// allocated on first use and kept until the process ends.  NULL on failure.
// Work that runs concurrently on two streams takes two slots.
Workspace* workspace_for(int res, int slot = 0);

}  // namespace silt
