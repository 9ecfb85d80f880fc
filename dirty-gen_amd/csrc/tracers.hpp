// Foam and stirred-silt tracers of the FluidCore window (tracers.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "core.hpp"

namespace silt {

// Inclusive cell bounds (i1 < 0: empty).
struct SpanBox {
  int i0, i1, j0, j1;
};

// Device-side control of one tracer pass: the live silt and foam and their
// bounds at the start, the visit box, the wet count, the silt's bounds as it
// moves, and whether the pass runs (the reference's early returns).
struct TracerState {
  int live_silt, live_foam;
  SpanBox foam_span, silt_span, silt_now;
  SpanBox box;
  int nwet;
  int run;
};

// The window's tracer planes: foam fraction and stirred silt (m), the live
// foam and this pass's wet and shore cells, 2 planes of scratch, and the
// FP64 accounts [0] stirred and [1] settled silt volume (m^3).
struct TracerPlanes {
  float* foam;
  float* turbid;
  uint8_t* foam_live;
  uint8_t* wet;
  uint8_t* shore;
  float* scratch;
  double* acc;
};

struct TracerPushes {
  int n;
  CorePush p[kMaxPush];
};

// step_tracers (src/fluid_core.rs:6210-6239) with the step's pushes; wake:
// a push at 0.25 m/s or slipping 5 %; push_box: the union of the pushes' 8 m
// reach in window cells (empty if none is on the window); foam_decay =
// exp(-dt / 1.8), silt_keep = exp(-dt / 4).
void tracer_step(const TracerPlanes& tr, const CorePlanes& c, TracerState* s, const TracerPushes& tp, int wake,
                 const SpanBox& push_box, float dt, float foam_decay, float silt_keep, float ox, float oz,
                 hipStream_t st);
// note_surface_foam (:6526-6538) with the last step's wake and reach.
void tracer_note_foam(const TracerPlanes& tr, const CorePlanes& c, TracerState* s, int wake, const SpanBox& push_box,
                      float dt, hipStream_t st);

}  // namespace silt
