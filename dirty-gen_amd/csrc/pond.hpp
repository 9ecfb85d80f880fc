// Ponds in tire cuts on the FluidCore window (pond.hip): pond_ruts and
// seat_rut_water with shove_rut_slurry (src/fluid_core.rs:5509-5792).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "core.hpp"

namespace silt {

// Scratch of the pond passes, allocated once per core (kWindowRes^2 cells,
// summed-area planes of (kWindowRes + 1)^2).
struct PondScratch {
  float* run;    // 2 planes: row running sums of the still mark and the still free surface
  float* pre;    // 2 summed-area planes
  float* f0;     // pond: the cut's extra depth; seat: the pressed ground
  float* f1;     // pond: the donor's room;      seat: the rut cap's dropped volume
  float* f2;     // seat: the bank cap's dropped volume
  float* f3;     // seat: slurry scraped off a cut (m)
  float* f4;     // seat: the share each ring cell takes
  float* old_w;  // water and slurry before the pass (the FP64 ledger line)
  float* old_s;
  uint8_t* m0;   // pond: still cells; seat: floor cells
  uint8_t* m1;   // seat: cut cells
  int8_t* ring;  // seat: the ring a cut's slurry goes to (0: none)
};

// What a pass returns: pond_ruts' moved volume (m^3); seat_rut_water's pond
// surface and whether it found a floor (1), found none (0) or was skipped on
// a dry uncut window (-1).
struct PondResult {
  float moved;
  float eta;
  int has_floor;
  int pad;
};

// pond_ruts (src/fluid_core.rs:5509-5611) on the window's water.  handoff[0]
// gains the FP64 change of the window's water volume (m^3): what the f32
// sums and the donors' max(0) leave of the exact transfer.
void core_pond_ruts(const CorePlanes& c, const PondScratch& s, PondResult* r, double* handoff, hipStream_t st);

// seat_rut_water (:5622-5736) and shove_rut_slurry (:5740-5792) with the
// classify closure in s.f0 (pressed ground) and s.m1 (cut flags).  gate:
// skip a window whose coarse sheet is dry and holds no cut (the caller in
// src/world.rs:5539-5561).  handoff[0..1] gain the FP64 change of the water
// and slurry volumes (m^3): the f32 rounding of the moves and the water that
// leaves the sheet when the floor has no room for it.
void core_seat_rut_water(const CorePlanes& c, const PondScratch& s, int gate, PondResult* r, double* handoff,
                         hipStream_t st);

// The classify closure of ThroughDrive's seat_rut_sheet (src/world.rs:
// 5546-5557) from the ground on the window grid: pressed = bed + placed
// (placed may be null), cut where initial height - bed > 12 mm.
void core_seat_classify(const float* soil_bed, const float* soil_placed, const float* soil_init, const PondScratch& s,
                        hipStream_t st);

}  // namespace silt
