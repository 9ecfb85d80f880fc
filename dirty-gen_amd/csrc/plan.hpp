// Host-side planning of one sheet step: which work boxes, how many substeps,
// which confinement schedule.  The rules are the reference's
// (step_sheet_scratch, src/fluid_core.rs:1659-1831).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "sheet_device.hpp"

namespace silt {

struct StepPlan {
  std::vector<Box> boxes;  // empty: the step is a no-op
  int nsub = 0;
  bool pace3 = false;      // one pressure pass, then nsub transport slices
  bool fine_open = false;
};

// sheet_work_boxes (src/fluid_core.rs:1325-1447).
std::vector<Box> sheet_work_boxes(const float* h, int res, float eps, bool allow_split);
// The split step of sheet_work_boxes given the bounding boxes of the wet
// bodies: each grown by 6 (clamped to the grid), touching boxes merged; one
// box, or more than 8, gives halo3 (the wet bounding box grown by 3).
std::vector<Box> boxes_from_bodies(std::vector<Box> bodies, int res, const Box& halo3);
// awake_work_boxes (src/fluid_core.rs:1537-1657).
std::vector<Box> awake_work_boxes(const float* h, const float* u, const float* v, const float* bed, int res,
                                  float cell, float eps);
// The boxes of awake_work_boxes from its 8 x 8 tiles' awake flags
// (row-major, ceil(res / 8) per side): grown by one tile, 4-connected.
std::vector<Box> awake_boxes_from_hot(const std::vector<uint8_t>& hot, int res);

// Whole plan from host copies of the planes.  Returns an empty plan when the
// reference's early returns fire (dry sheet, dt <= 0).
StepPlan plan_step(const float* h, const float* u, const float* v, const float* bed, int res, float cell, float dt,
                   int wind, float wx, float wz, float eps, float sub_cap, int pace, bool walled = false);

// True when the reference sweeps confinement from a snapshot (Jacobi) on
// this box: the 2 cm gravity sheet, or a box of >= 48 x 32 cells when row
// splits are allowed (force_gravity_open :2221-2234, force_wind_open :2649).
bool reference_uses_jacobi(const Box& b, float cell, int wind, bool row_split);

}  // namespace silt
