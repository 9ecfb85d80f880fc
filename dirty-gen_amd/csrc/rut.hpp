// Bekker rut coupling (rut.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "sheet_device.hpp"
#include "silt_rut.h"

namespace silt {

// One prepared contact: host-side scalars of stamp_wheel's tread branch.
struct RutContact {
  float cx, cz, hub_y, fx, fz, rx, rz, radius, sinkage;
  float forced, want, e_fast, e_slow, tread_core, tread_edge, along_reach;
  int ilo, ihi, jlo, jhi;  // footprint, [lo, hi)
  int loaded, skip;
};

float bekker_load(float sinkage, float radius, float width, float k_eq, float n);
float sinkage_with(float load, float radius, float width, float k_eq, float n);
float rut_depth_from_force(float normal, float push, float radius, float width, float k_eq, float n);
float contact_length(float radius, float sinkage);

int prepare_contacts(const silt_tire_contact* in, size_t n, int res, float cell, float ox, float oz, float dt,
                     std::vector<RutContact>* out);
// The same into out[0..n) (no allocation).
int prepare_contacts_into(const silt_tire_contact* in, size_t n, int res, float cell, float ox, float oz, float dt,
                          RutContact* out);
// Copies nc prepared contacts (host) to d_scratch and enqueues the stamp
// kernel over them (the launch of rut_stamp).
int rut_stamp_prepared(float* bed, float* placed, const float* initial, int res, float cell, float ox, float oz,
                       const RutContact* host, int nc, RutContact* d_scratch, hipStream_t st, int per_group = 0,
                       float* sheet_bed = nullptr, float* rut = nullptr);
// Stamps the contacts in order on device planes; d_scratch holds >= n
// contacts.  per_group > 0: consecutive groups of that many contacts whose
// footprints do not overlap each other are stamped in parallel (in order
// within a group); 0: all in one sequence.  sheet_bed / rut set: the
// footprints' sheet bed (bed + placed) and rut (max(initial - bed, 0)) are
// refreshed too, so the caller needs no full-plane pass.
int rut_stamp(float* bed, float* placed, const float* initial, int res, float cell, float ox, float oz,
              const silt_tire_contact* contacts, size_t n, float dt, RutContact* d_scratch, hipStream_t st,
              int per_group = 0, float* sheet_bed = nullptr, float* rut = nullptr);

}  // namespace silt
