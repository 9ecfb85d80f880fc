// Device structures of the two-layer FluidCore (api_core.cpp, core_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sheet_device.hpp"

namespace silt {

constexpr int kWindowRes = 200;     // WINDOW_RES, src/fluid_core.rs:72
constexpr int kFineRes = 160;       // FINE_RES, :76
constexpr float kCoarseCell = 0.20f;  // COARSE_CELL_M, :70
constexpr int kMaxPush = 64;
// Cell areas of the FP64 ledger.  A coarse cell is exactly 10 x 10 patch
// cells, so the patch's area is the coarse one over 100 (the FP32 cell sides
// 0.2 and 0.02 would make a refine or restrict change the volume by 7e-8).
constexpr double kCoarseArea = (double)kCoarseCell * (double)kCoarseCell;
constexpr double kFineArea = kCoarseArea / 100.0;
constexpr unsigned long long kEmptyKey = 0xFFFFFFFFFFFFFFFFull;
constexpr unsigned long long kTombKey = 0xFFFFFFFFFFFFFFFEull;

struct GridGeo {
  int res;
  float cell;
  float ox, oz;  // world origin of cell (0, 0)'s corner
};

// TirePush (src/fluid_core.rs:357-368).
struct CorePush {
  float x, z, fx, fz, radius, width, speed, slip;
};

struct PushList {
  int n;
  CorePush p[kMaxPush];
};

// A group of pushes (p0 .. p0 + np of a push array) and the cells it shoves.
struct PushBox {
  int p0, np;
  int i0, j0, w, h;
};

// SprayEvent (src/fluid_core.rs:371-382).
struct SprayEvent {
  float x, y, z, vx, vy, vz, volume;
  int kind;
};

struct CorePlanes {
  float *water, *slurry, *uw, *vw, *us, *vs, *bed, *moisture;
  uint8_t* grain;
};

struct FinePlanes {
  float *water, *slurry, *uw, *vw, *us, *vs, *bed, *bed_s, *moisture;
  uint8_t* grain;
};

// Coarse column / row of every fine column / row, -1 off the window
// (fine_coarse_axes, src/fluid_core.rs:9518-9534).
struct FineAxes {
  short col[kFineRes];
  short row[kFineRes];
};

// The coarse cells the patch covers and, for each coarse column / row, its
// first fine column / row and how many there are.
struct CoarseSpans {
  int ci0, cj0, ncol, nrow;
  short col_start[24], col_count[24], row_start[24], row_count[24];
};

struct NearColumn {
  int cidx, kx, kz, i0, j0, inner;
};

// Near-field residual tiles in an open-addressing table on the device.
struct NearTable {
  unsigned long long* keys;  // (kx, kz) packed, kEmptyKey / kTombKey
  float* data;               // 400 floats per slot: w, s, uw, vw residuals
  float* peak;
  float* idle;
  int* fresh;
  unsigned* count;
  unsigned cap;
};

void core_push_on(const GridGeo& g, float* water, float* slurry, float* uw, float* vw, float* us, float* vs,
                  const PushList& pl, int i0, int j0, int w, int h, hipStream_t st);
// push_on of n groups whose boxes do not overlap, in one launch (boxes and
// pushes in device memory; max_w / max_h bound the boxes).
void core_push_on_many(const GridGeo& g, float* water, float* slurry, float* uw, float* vw, float* us, float* vs,
                       const PushBox* boxes, const CorePush* pushes, int n, int max_w, int max_h, hipStream_t st);
void core_peel(const GridGeo& g, float* water, const float* uw, const float* vw, const PushList& pl, int i0, int j0,
               int w, int h, const uint8_t* hold, const GridGeo& cg, bool skip_held, SprayEvent* ev, unsigned* nev,
               unsigned cap, double* vol, double area, hipStream_t st);
void core_soften(float* u, float* v, int res, hipStream_t st);
void core_nu(float* nu, const uint8_t* grain, const float* moisture, int n, hipStream_t st);
void core_fill(float* p, float val, int n, hipStream_t st);
void core_fill_bytes(uint8_t* p, uint8_t val, int n, hipStream_t st);
void core_refine_copy(const CorePlanes& c, const FinePlanes& f, const uint8_t* hold, const FineAxes& ax, int ci0,
                      int cj0, int cw, int ch, hipStream_t st);
// frozen[0..1]: FP64 water / slurry volume parked off the window.
// handoff[0..1]: FP64 volume the coarse cells gained beyond the patch volume
// handed to them (the FP32 sums' rounding).
void core_restrict(const CorePlanes& c, const FinePlanes& f, const uint8_t* hold, const CoarseSpans& sp,
                   const FineAxes& ax, double* frozen, double* handoff, hipStream_t st);
void core_near_capture(const FinePlanes& f, const NearTable& t, const NearColumn* cols, int ncols, hipStream_t st);
// handoff[0..1]: FP64 volume the lay-down added to the water / slurry (the
// stored detail's FP32 mean and the max(0) clamp).
void core_near_apply(const FinePlanes& f, const NearTable& t, const NearColumn* cols, int ncols, double* handoff,
                     hipStream_t st);
void core_near_fade(const NearTable& t, float dt, hipStream_t st);
void core_near_clear(const NearTable& t, hipStream_t st);
void core_near_rehash(const NearTable& from, const NearTable& to, hipStream_t st);
void core_ground(const float* soil_bed, const float* soil_placed, const float* soil_init, float* sheet_bed, float* rut,
                 int n, hipStream_t st);
// susp / turbid: the suspended and stirred solids (m) behind flow's
// turbidity channel, either may be null.
void core_publish(const CorePlanes& c, const float* soil_bed, const float* soil_placed, const float* soil_rut,
                  const float* susp, const float* turbid, float* ground_rut4, float* film2, float* flow4,
                  hipStream_t st);
// One hydro solid as the wall mask tests it (solid_mask, src/fluid_core.rs:
// 7288-7362): the rectangle's centre, axes (sin / cos of the yaw, from the
// host), half sizes, the heights between which it is a wall, the cell range
// its bounding circle covers, and the mask value (2 wall, 1 contact).
struct SolidDesc {
  float x, z, fwd_x, fwd_z, half_length, half_width, bottom_y, top_y;
  int i0, i1, j0, j1;
  int value;
  int pad[3];
};
constexpr int kMaxSolids = 64;

// The wall mask of one grid from its water and bed: walls first, then
// contacts (descs in that order).
void core_solid_mask(const float* water, const float* bed, int res, float cell, float ox, float oz,
                     const SolidDesc* descs, int n, uint8_t* mask, hipStream_t st);
// scour_contacts (src/fluid_core.rs:7364-7421) on the window: bed into
// slurry beside contacts.  acc[0]: FP64 eroded volume (m^3), acc[1]: the
// deepest drop (float bits); source: the slurry volume the ledger books.
void core_scour(const CorePlanes& c, const uint8_t* mask, float* bed_drop, float dt, double* eroded,
                unsigned* hole_bits, double* source, hipStream_t st);

// moisture (max of the sheet's and the soil's wetness), wet_edge and
// foam_crest (foam tracer, placed + film excess) maps; wetness / placed /
// foam may be null.
void core_publish_surface(const CorePlanes& c, const float* wetness, const float* placed, const float* foam, int res,
                          float* moisture1, float* wet_edge1, float* foam2, hipStream_t st);

}  // namespace silt
