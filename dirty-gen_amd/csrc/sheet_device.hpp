// Shared declarations of the sheet-step kernels (device laws, launch
// arguments, plane bundle) for the exact and the fused paths.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace silt {

constexpr float kGravity = 9.81f;          // GRAVITY, src/fluid_core.rs:110
constexpr float kFineCell = 0.02f;         // FINE_CELL_M, :74
constexpr float kFilmSpeedCap = 1.6f;      // FILM_SPEED_CAP_MPS, :10003

// Inclusive work box (i0, i1, j0, j1) as in sheet_work_boxes.
struct Box {
  int i0, i1, j0, j1;
};

// Running FP64 ledger lines of one layer, metres of depth summed over cells:
// depth the dry cut deleted, and the transport residue (FP32 rounding and the
// max(0) clamp beyond the exchanged fluxes).
struct SheetTotals {
  double cut;
  double residue;
};

// Everything one substep of one layer needs besides the planes.
struct StepArgs {
  int res;
  float cell;
  float dt;     // substep dt (the forces' dt; transport slices get their own)
  int wind;     // 0 gravity drive, 1 wind drive
  float wx, wz;
  float vort;   // confinement strength
  float eps;    // dry threshold
  int project;  // divergence correction on (coarse gravity sheet)
  float lim;    // speed cap
  float blend;  // wind blend 1 - exp(-1.6 dt)
  int banks;    // dry-bank test in transport
  int gate_uniform;
  // Viscosity-derived scalars when every cell has the same viscosity.
  int nu_uniform;
  float nu0, mob0, ov0, decay0;
  // Ledger lines the transport and the dry cut add to, or null.
  SheetTotals* book;
};

// Device planes of one layer plus the scratch one substep uses.
struct SheetPlanes {
  float* h;
  float* u;
  float* v;
  const float* bed;
  const float* nu;
  const uint8_t* gate;
  // scratch
  float* eta;
  float* tg;
  float* ts;
  float* p;
  float* su;
  float* sv;
  float* bsrc;
  float* h1;
  // Wall mask (nonzero = wall) of the walled sheet, or null.
  const uint8_t* solid = nullptr;
};

// Rust f32::clamp: NaN passes through.
__host__ __device__ inline float clampf(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}

// x / d for a float divisor d, correctly rounded without a division, given
// rd = RN64(1 / d): the double product x * rd is within 2^-52 (relative) of
// the quotient, and a quotient of two floats lies at least 2^-49 (relative)
// from every float rounding boundary (x = m d for a 25-bit midpoint m would
// need x to carry 25 significant bits), so one rounding to float gives the
// IEEE quotient for every x: zeros, subnormals, infinities and NaNs too.
// Three instructions against the dozen of an IEEE division.
__host__ __device__ inline float div_by(float x, double rd) { return (float)((double)x * rd); }

// sheet_mobility_for, src/fluid_core.rs:169-180.
__host__ __device__ inline float sheet_mobility(float nu) {
  if (nu < 0.0f) return -nu;
  float boost = 0.0f;
  if (nu < 12.0f) {
    float t = clampf(div_by(12.0f - fmaxf(nu, 0.0f), 1.0 / 12.0), 0.0f, 1.0f);
    boost = t * t * 1.6f;
  }
  return 0.042f * (1.0f + boost);
}

// sheet_damp, src/fluid_core.rs:184-192.
__host__ __device__ inline float sheet_damp(float nu) {
  float rel = clampf(div_by(fmaxf(nu, 0.05f), 1.0 / 231.0), 0.002f, 12.0f);
  if (rel >= 0.42f) return 1.70f * powf(rel, 0.30f);
  float at_silt = 1.70f * powf(0.42f, 0.30f);
  float t = clampf(div_by(rel, 1.0 / (double)0.42f), 0.0f, 1.0f);
  return at_silt * powf(t, 1.15f);
}

// dynamic_viscosity, src/fluid_core.rs:150-162 (grain codes of Grain).
__host__ __device__ inline float dynamic_viscosity(int grain, float moisture) {
  float x = clampf((clampf(moisture, 0.0f, 1.0f) - 0.22f) / 0.70f, 0.0f, 1.0f);
  float thin = x * x * (3.0f - 2.0f * x);
  float stiff, fluid;
  switch (grain) {
    case 1: stiff = 220.0f; fluid = 4.0f; break;   // Silt
    case 2: stiff = 28.0f; fluid = 0.55f; break;   // Sand
    case 3: stiff = 0.9f; fluid = 0.9f; break;     // Water
    case 4: stiff = 1.5f; fluid = 0.08f; break;    // Other
    default: stiff = 420.0f; fluid = 0.35f; break; // Clay
  }
  return stiff + (fluid - stiff) * thin;
}

// overload of the divergence source, src/fluid_core.rs:1999.
__host__ __device__ inline float overload_of(float nu) { return 1.0f + clampf(div_by(nu, 1.0 / 231.0), 0.0f, 6.0f); }

// finite_speed, src/fluid_core.rs:3473-3479.
__host__ __device__ inline float finite_speed(float v, float lim) {
  if (!isfinite(v)) return 0.0f;
  return clampf(v, -lim, lim);
}

// Exact path (sheet_exact.hip).
void exact_substep(const SheetPlanes& s, const Box& b, const StepArgs& a, bool do_flux, bool gauss_seidel,
                   hipStream_t st);
void exact_transport(const SheetPlanes& s, const Box& b, const StepArgs& a, float dt, hipStream_t st);
void exact_dry_cut(const SheetPlanes& s, const Box& b, const StepArgs& a, hipStream_t st);
// One substep of one box beside the walls s.solid marks (par: the parallel
// kicks, then the confinement sweep).
void exact_substep_walled(const SheetPlanes& s, const Box& b, const StepArgs& a, bool par, hipStream_t st);
// displace_occupied over a whole grid (share, mode: scratch planes).
void displace_occupied(float* h, float* u, float* v, const uint8_t* solid, int res, float* share, float* mode,
                       SheetTotals* book, hipStream_t st);

}  // namespace silt
