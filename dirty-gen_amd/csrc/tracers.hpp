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

// The wheel-lifted suspended solids (src/fluid_core.rs:8750-9003): metres of
// solids in the water column; scratch per cell: a donor's give, its
// concentration, what its receiver accepted, what a receiver took, and the
// donor's direction; the FP64 accounts [0] entrained from the bed,
// [1] settled back (m^3).
struct SuspPlanes {
  float* susp;
  float* give;
  float* conc;
  float* accept;
  float* taken;
  int8_t* dir;
  double* acc;
};

// entrain_wake (:8758-8808) of the pushes, in push order.
void susp_entrain(const SuspPlanes& sp, const float* water, const TracerPushes& tp, float dt, float ox, float oz,
                  hipStream_t st);
// step_suspended / advect_suspended (:8750-9003) on a sheet that is not
// surface driven (no waves, no river reach: the plume rides v only, no eddy
// mixing).
void susp_step(const SuspPlanes& sp, const float* water, const float* vw, float dt, hipStream_t st);

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
