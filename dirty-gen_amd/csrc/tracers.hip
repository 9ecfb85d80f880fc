// Foam and stirred-silt tracers on the 20 cm window (step_tracers and
// note_surface_foam, src/fluid_core.rs:6210-6750, laws :9676-9880).
//
// The reference walks index lists of wet and shore cells and bails out early
// when nothing is live.  Here every list is a per-cell flag plane and every
// early return a flag in TracerState that the kernels read, so a step is a
// fixed sequence of launches with no host wait: scan (live silt, live foam,
// their bounds) -> visit box -> wet and shore flags -> foam advection ->
// silt advection -> settling -> tire, shore and Froude production -> dry
// cleanup -> live foam.  Each cell's updates keep the reference's order
// (tires in push order, then the shore, then Froude foam), so foam and silt
// equal the FP32 restatement value for value; the stirred and settled
// accounts, which the reference sums in f32 in list order, are FP64 sums.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tracers.hpp"

namespace silt {

namespace {

constexpr int R = kWindowRes;
constexpr float kFoamEq = 0.85f;     // FOAM_EQ, src/fluid_core.rs:115
constexpr float kFoamTau = 0.35f;    // FOAM_TAU_S, :117
constexpr float kSiltCap = 0.028f;   // SILT_COLUMN_CAP_M, :124

__device__ inline bool fin(float x) { return isfinite(x); }
__device__ inline float clampd(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

// f32::hypot as the reference's libm computes it: the sum of squares in
// double (exact products), one double square root, rounded to float.
__device__ inline float hypot_ref(float x, float y) {
  if (isinf(x) || isinf(y)) return INFINITY;
  return (float)__dsqrt_rn((double)x * (double)x + (double)y * (double)y);
}

__device__ inline float smooth_step(float e0, float e1, float x) {
  const float t = clampd((x - e0) / (e1 - e0), 0.0f, 1.0f);
  return t * t * (3.0f - 2.0f * t);
}

// breaking_foam (:9676-9686)
__device__ inline float breaking_foam(float speed, float h) {
  if (!fin(speed) || !fin(h) || h < 0.004f || speed <= 0.05f) return 0.0f;
  const float froude = speed / sqrtf(kGravity * h);
  return smooth_step(0.85f, 1.45f, froude);
}

// add_foam (:6400-6419)
__device__ inline float add_foam(float foam, float prod, float dt) {
  if (prod <= 0.0f || !fin(prod) || !(dt > 0.0f)) return foam;
  const float cur = fin(foam) ? clampd(foam, 0.0f, kFoamEq) : 0.0f;
  const float room = kFoamEq - cur;
  if (room <= 1.0e-5f) return foam;
  const float grown = cur + room * prod * fminf(dt / kFoamTau, 1.0f);
  return fin(grown) ? fminf(grown, kFoamEq) : 0.0f;
}

// add_silt (:6704-6715): returns the depth added.
__device__ inline float add_silt(float* t, float dh) {
  if (dh <= 0.0f || !fin(dh)) return 0.0f;
  const float room = fmaxf(kSiltCap - fmaxf(*t, 0.0f), 0.0f);
  const float add = fminf(dh, room);
  if (add <= 0.0f) return 0.0f;
  *t = *t + add;
  return add;
}

__device__ inline void wave_add(double* p, double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0 && v != 0.0) atomicAdd(p, v);
}

__device__ inline int cell_index() { return (int)(blockIdx.x * blockDim.x + threadIdx.x); }

__global__ void k_tr_reset(TracerState* s) {
  if (threadIdx.x != 0) return;
  s->live_silt = s->live_foam = 0;
  for (SpanBox* b : {&s->foam_span, &s->silt_span, &s->silt_now}) {
    b->i0 = b->j0 = R;
    b->i1 = b->j1 = -1;
  }
  s->nwet = 0;
  s->run = 0;
}

__device__ inline void span_add(SpanBox* b, int i, int j) {
  atomicMin(&b->i0, i);
  atomicMax(&b->i1, i);
  atomicMin(&b->j0, j);
  atomicMax(&b->j1, j);
}

// The live silt and live foam and their bounds (step_tracers :6220-6222,
// turbid_span :9766-9794, tracer_visit_box :6292-6337).
__global__ void k_tr_scan(const float* __restrict__ turbid, const uint8_t* __restrict__ foam_live, TracerState* s) {
  const int idx = cell_index();
  if (idx >= R * R) return;
  const int i = idx % R, j = idx / R;
  const float t = turbid[idx];
  if (t > 1.0e-8f) {
    s->live_silt = 1;
    span_add(&s->silt_span, i, j);
  }
  if (foam_live[idx]) {
    s->live_foam = 1;
    span_add(&s->foam_span, i, j);
  }
}

// The silt bounds of the current planes only (the spans the advection,
// settling and cleanup walk are recomputed as the silt moves).
__global__ void k_tr_span_reset(SpanBox* b) {
  if (threadIdx.x != 0) return;
  b->i0 = b->j0 = R;
  b->i1 = b->j1 = -1;
}
__global__ void k_tr_span(const float* __restrict__ turbid, SpanBox* b) {
  const int idx = cell_index();
  if (idx >= R * R) return;
  if (turbid[idx] > 1.0e-8f) span_add(b, idx % R, idx / R);
}

// turbid_span's growth by one cell; false when no cell holds silt.
__device__ inline bool grown(const SpanBox& b, int* i0, int* i1, int* j0, int* j1) {
  if (b.i1 < 0) return false;
  *i0 = b.i0 > 0 ? b.i0 - 1 : 0;
  *i1 = b.i1 + 1 < R - 1 ? b.i1 + 1 : R - 1;
  *j0 = b.j0 > 0 ? b.j0 - 1 : 0;
  *j1 = b.j1 + 1 < R - 1 ? b.j1 + 1 : R - 1;
  return true;
}

// The visit box (collect_tracer_wet :6246-6290): the window when a tire
// wakes it, else the live foam, the silt and the tires' reach, plus 2.
// note: note_surface_foam skips the step's early return.
__global__ void k_tr_box(TracerState* s, int wake, SpanBox pushes, int note) {
  if (threadIdx.x != 0) return;
  if (!note && !wake && !s->live_silt && !s->live_foam) {
    s->run = 0;
    return;
  }
  if (wake) {
    s->box = SpanBox{0, R - 1, 0, R - 1};
    s->run = 1;
    return;
  }
  int a = R, b = 0, c = R, d = 0;
  bool any = false;
  if (s->foam_span.i1 >= 0) {
    a = min(a, s->foam_span.i0);
    b = max(b, s->foam_span.i1);
    c = min(c, s->foam_span.j0);
    d = max(d, s->foam_span.j1);
    any = true;
  }
  int ti0, ti1, tj0, tj1;
  if (grown(s->silt_span, &ti0, &ti1, &tj0, &tj1)) {
    a = min(a, ti0);
    b = max(b, ti1);
    c = min(c, tj0);
    d = max(d, tj1);
    any = true;
  }
  if (pushes.i1 >= 0) {
    a = min(a, pushes.i0);
    b = max(b, pushes.i1);
    c = min(c, pushes.j0);
    d = max(d, pushes.j1);
    any = true;
  }
  if (!any) {
    s->run = 0;
    return;
  }
  s->box = SpanBox{a >= 2 ? a - 2 : 0, b + 2 < R - 1 ? b + 2 : R - 1, c >= 2 ? c - 2 : 0, d + 2 < R - 1 ? d + 2 : R - 1};
  s->run = 1;
}

// shore_wet_cell (:9752-9764)
__device__ inline bool shore_wet(const float* water, int i, int j) {
  if (i == 0 || j == 0 || i + 1 >= R || j + 1 >= R) return false;
  const int idx = j * R + i;
  const int nb[4] = {idx - 1, idx + 1, idx - R, idx + R};
  for (int q = 0; q < 4; q++) {
    const float nh = water[nb[q]];
    if (fin(nh) && nh < 0.004f) return true;
  }
  return false;
}

__global__ void k_tr_collect(const float* __restrict__ water, const float* __restrict__ slurry, TracerState* s,
                             uint8_t* __restrict__ wet, uint8_t* __restrict__ shore) {
  const int idx = cell_index();
  const bool in = idx < R * R;
  bool w = false;
  if (in) {
    const int i = idx % R, j = idx / R;
    bool sh = false;
    if (s->run && i >= s->box.i0 && i <= s->box.i1 && j >= s->box.j0 && j <= s->box.j1) {
      const float h = water[idx] + slurry[idx];
      w = fin(h) && !(h < 0.004f);
      sh = w && shore_wet(water, i, j);
    }
    wet[idx] = w;
    shore[idx] = sh;
  }
  const uint64_t m = __builtin_amdgcn_ballot_w64(w);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(&s->nwet, (int)__builtin_popcountll(m));
}

// The step proceeds past collect_tracer_wet (:6225-6227).
__device__ inline bool proceed(const TracerState* s) { return s->run && (s->nwet > 0 || s->live_silt); }

// sample_foam (:9693-9714)
__device__ inline float sample_foam(const float* src, float x, float z) {
  if (!fin(x) || !fin(z) || x < 0.0f || z < 0.0f) return 0.0f;
  const float fx = floorf(x), fz = floorf(z);
  if (!(fx < (float)R) || !(fz < (float)R)) return 0.0f;
  const int i0 = (int)fx, j0 = (int)fz;
  const int i1 = i0 + 1 < R ? i0 + 1 : R - 1, j1 = j0 + 1 < R ? j0 + 1 : R - 1;
  const float tx = clampd(x - (float)i0, 0.0f, 1.0f), tz = clampd(z - (float)j0, 0.0f, 1.0f);
  const float a = src[j0 * R + i0], b = src[j0 * R + i1], c = src[j1 * R + i0], d = src[j1 * R + i1];
  const float v = a + (b - a) * tx + (c - a) * tz + (a - b - c + d) * tx * tz;
  return fin(v) ? clampd(v, 0.0f, 1.0f) : 0.0f;
}

// advect_foam (:6352-6398), the sampling half.
__global__ void k_foam_sample(const float* __restrict__ foam, const float* __restrict__ uw,
                              const float* __restrict__ vw, const uint8_t* __restrict__ wet, const TracerState* s,
                              float dt, float* __restrict__ sample) {
  const int idx = cell_index();
  if (idx >= R * R || !proceed(s) || !s->live_foam || !wet[idx]) return;
  const float u = uw[idx], v = vw[idx];
  float out = 0.0f;
  if (fin(u) && fin(v)) {
    const float x = (float)(idx % R) - u * dt / kCoarseCell;
    const float z = (float)(idx / R) - v * dt / kCoarseCell;
    out = sample_foam(foam, x, z);
  }
  sample[idx] = out;
}

// The writing half: live and wet cells start from 0, wet ones take the
// decayed sample.
__global__ void k_foam_write(float* __restrict__ foam, const uint8_t* __restrict__ wet,
                             const uint8_t* __restrict__ live, const TracerState* s, float decay,
                             const float* __restrict__ sample) {
  const int idx = cell_index();
  if (idx >= R * R || !proceed(s) || !s->live_foam) return;
  if (!(live[idx] || wet[idx])) return;
  float f = 0.0f;
  if (wet[idx]) {
    const float v = sample[idx] * decay;
    if (v > 1.0e-4f && fin(v)) f = fminf(v, kFoamEq);
  }
  foam[idx] = f;
}

// flux_edge (:9796-9824) out of a donor whose book already holds `taken`.
__device__ inline float edge_out(float column, float taken, float vel, float dt) {
  if (!fin(vel) || fabsf(vel) < 1.0e-5f) return 0.0f;
  const float room = column - taken;
  if (room <= 1.0e-8f) return 0.0f;
  const float frac = fminf(fabsf(vel) * dt / kCoarseCell, 0.45f);
  const float dh = fminf(column * frac, room);
  if (dh <= 1.0e-8f || !fin(dh)) return 0.0f;
  return dh;
}

// The four outflows of cell (i, j) in flux_thickness's edge order (its
// left, right, upper and lower faces, each only toward that neighbour and
// only where the face lies in the span), each with the room its earlier
// ones left.
struct Out4 {
  float l, r, u, d;
};
__device__ inline Out4 outflows(const float* t, const float* uw, const float* vw, int i, int j, int i0, int i1,
                                int j0, int j1, float dt) {
  Out4 o{0.0f, 0.0f, 0.0f, 0.0f};
  if (i < i0 || i > i1 || j < j0 || j > j1) return o;
  const int idx = j * R + i;
  const float tv = t[idx];
  const float column = fin(tv) ? tv : 0.0f;
  const int xe = i1 < R - 1 ? i1 : R - 1;  // x faces (i, i + 1) for i < xe
  const int ze = j1 < R - 1 ? j1 : R - 1;  // z faces (j, j + 1) for j < ze
  float taken = 0.0f;
  if (i - 1 >= i0 && i - 1 < xe) {
    const float vel = 0.5f * (uw[idx - 1] + uw[idx]);
    if (!(vel > 0.0f)) {
      o.l = edge_out(column, taken, vel, dt);
      taken = taken + o.l;
    }
  }
  if (i < xe) {
    const float vel = 0.5f * (uw[idx] + uw[idx + 1]);
    if (vel > 0.0f) {
      o.r = edge_out(column, taken, vel, dt);
      taken = taken + o.r;
    }
  }
  if (j - 1 >= j0 && j - 1 < ze) {
    const float vel = 0.5f * (vw[idx - R] + vw[idx]);
    if (!(vel > 0.0f)) {
      o.u = edge_out(column, taken, vel, dt);
      taken = taken + o.u;
    }
  }
  if (j < ze) {
    const float vel = 0.5f * (vw[idx] + vw[idx + R]);
    if (vel > 0.0f) o.d = edge_out(column, taken, vel, dt);
  }
  return o;
}

// advect_turbid (:6421-6471): the span's new columns into `next`.
__global__ void k_silt_advect(const float* __restrict__ t, const float* __restrict__ uw, const float* __restrict__ vw,
                              const TracerState* s, float dt, float* __restrict__ next) {
  const int idx = cell_index();
  if (idx >= R * R || !proceed(s)) return;
  int i0, i1, j0, j1;
  if (!grown(s->silt_span, &i0, &i1, &j0, &j1)) return;
  const int i = idx % R, j = idx / R;
  if (i < i0 || i > i1 || j < j0 || j > j1) return;
  const Out4 me = outflows(t, uw, vw, i, j, i0, i1, j0, j1, dt);
  const float rem = ((me.l + me.r) + me.u) + me.d;
  float add = 0.0f;
  if (i > 0) add = add + outflows(t, uw, vw, i - 1, j, i0, i1, j0, j1, dt).r;
  if (i + 1 < R) add = add + outflows(t, uw, vw, i + 1, j, i0, i1, j0, j1, dt).l;
  if (j > 0) add = add + outflows(t, uw, vw, i, j - 1, i0, i1, j0, j1, dt).d;
  if (j + 1 < R) add = add + outflows(t, uw, vw, i, j + 1, i0, i1, j0, j1, dt).u;
  const float tv = t[idx];
  float out;
  if (rem == 0.0f && add == 0.0f) {
    out = fin(tv) ? tv : 0.0f;
  } else {
    const float nx = tv - rem + add;
    out = fin(nx) ? fmaxf(nx, 0.0f) : 0.0f;
  }
  next[idx] = out;
}

__global__ void k_silt_commit(float* __restrict__ t, const TracerState* s, const float* __restrict__ next) {
  const int idx = cell_index();
  if (idx >= R * R || !proceed(s)) return;
  int i0, i1, j0, j1;
  if (!grown(s->silt_span, &i0, &i1, &j0, &j1)) return;
  const int i = idx % R, j = idx / R;
  if (i < i0 || i > i1 || j < j0 || j > j1) return;
  t[idx] = next[idx];
}

// settle_turbid (:6473-6509) over the span of the carried silt.
__global__ void k_silt_settle(float* __restrict__ t, const float* __restrict__ water, const float* __restrict__ slurry,
                              const TracerState* s, float keep, double* settled) {
  const int idx = cell_index();
  double landed = 0.0;
  int i0, i1, j0, j1;
  if (idx < R * R && proceed(s) && grown(s->silt_now, &i0, &i1, &j0, &j1)) {
    const int i = idx % R, j = idx / R;
    if (i >= i0 && i <= i1 && j >= j0 && j <= j1) {
      const float h = water[idx] + slurry[idx];
      const float tv = t[idx];
      const float column = fin(tv) ? fmaxf(tv, 0.0f) : 0.0f;
      if (column <= 0.0f) {
        t[idx] = 0.0f;
      } else if (!fin(h) || h < 1.0e-5f) {
        landed = (double)(column * (kCoarseCell * kCoarseCell));
        t[idx] = 0.0f;
      } else {
        const float stay = column * keep;
        landed = (double)((column - stay) * (kCoarseCell * kCoarseCell));
        t[idx] = stay;
      }
    }
  }
  wave_add(settled, landed);
}

// shore_cell (:9826-9874)
__device__ inline void shore_cell(const float* water, const float* uw, const float* vw, int idx, float* foam_out,
                                  float* rate_out) {
  *foam_out = 0.0f;
  *rate_out = 0.0f;
  const int i = idx % R, j = idx / R;
  if (i == 0 || j == 0 || i + 1 >= R || j + 1 >= R) return;
  const float h = water[idx];
  if (!fin(h) || h < 0.004f) return;
  float nx = 0.0f, nz = 0.0f;
  bool edge = false;
  const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
  for (int q = 0; q < 4; q++) {
    const float nh = water[(j + dj[q]) * R + i + di[q]];
    if (fin(nh) && nh < 0.004f) {
      edge = true;
      nx = nx + (float)di[q];
      nz = nz + (float)dj[q];
    }
  }
  if (!edge) return;
  const float u = uw[idx], v = vw[idx];
  if (!fin(u) || !fin(v)) return;
  const float film = hypot_ref(u, v);
  const float nlen = fmaxf(hypot_ref(nx, nz), 1.0e-4f);
  const float into = fmaxf((u * nx + v * nz) / nlen, 0.0f);
  const float strain = fabsf(uw[idx + 1] - uw[idx - 1]) + fabsf(vw[idx + R] - vw[idx - R]);
  const float shear_v = fmaxf(film, strain);
  if (shear_v < 0.2f && into < 0.2f) return;
  const float prod = fmaxf(breaking_foam(fmaxf(shear_v, into), h), smooth_step(0.25f, 1.1f, into));
  *foam_out = prod > 0.02f ? prod : 0.0f;
  if (h > 0.012f && h < 0.055f && (shear_v > 0.45f || into > 0.4f)) {
    const float tau = 1000.0f * 0.025f * shear_v * shear_v;
    if (tau > 3.0f || into > 0.45f) *rate_out = fminf(0.006f * (into + shear_v * 0.35f), 0.01f);
  }
}

// spawn_foam_and_silt (:6511-6524): each wet cell takes every tire's ring
// foam and patch silt in push order (spawn_tire_tracers :6602-6669), then
// its shore foam and silt (:6671-6689), then Froude foam (:6542-6561).
__global__ void k_tr_spawn(float* __restrict__ foam, float* __restrict__ t, const float* __restrict__ water,
                           const float* __restrict__ slurry, const float* __restrict__ uw,
                           const float* __restrict__ vw, const uint8_t* __restrict__ wet,
                           const uint8_t* __restrict__ shore, const TracerState* s, TracerPushes tp, float dt,
                           float ox, float oz, double* stirred) {
  const int idx = cell_index();
  double lifted = 0.0;
  if (idx < R * R && proceed(s) && wet[idx]) {
    const float cell = kCoarseCell, area = kCoarseCell * kCoarseCell;
    float f = foam[idx], tv = t[idx];
    const float hw = water[idx];
    const int i = idx % R, j = idx / R;
    const float x = ox + ((float)i + 0.5f) * cell;
    const float z = oz + ((float)j + 0.5f) * cell;
    for (int k = 0; k < tp.n; k++) {
      const CorePush& p = tp.p[k];
      const float speed = fin(p.speed) ? fmaxf(p.speed, 0.0f) : 0.0f;
      const float slip = fin(p.slip) ? fminf(fabsf(p.slip), 1.5f) : 0.0f;
      if (speed < 0.25f && slip < 0.05f) continue;
      if (!fin(hw) || hw < 0.004f) continue;
      const float radius = fmaxf(p.radius, 0.05f);
      const float len = fmaxf(sqrtf(p.fx * p.fx + p.fz * p.fz), 1.0e-4f);
      const float fx = p.fx / len, fz = p.fz / len;
      const float rx = fz, rz = -fx;
      const float reach = radius * 1.15f + cell;
      const float reach2 = reach * reach;
      const float slip_v = speed * slip;
      const float dx = x - p.x, dz = z - p.z;
      if (dx * dx + dz * dz > reach2) continue;
      const float along = dx * fx + dz * fz;
      const float lat = dx * rx + dz * rz;
      const float dist = hypot_ref(dx, dz);
      float film = hypot_ref(uw[idx], vw[idx]);
      film = fin(film) ? film : 0.0f;
      const float immerse = fminf(hw, radius * 0.9f);
      const float r_line = sqrtf(fmaxf(2.0f * radius * immerse - immerse * immerse, 0.0f));
      const float band = 1.0f - clampd(fabsf(dist - r_line) / (cell * 1.35f), 0.0f, 1.0f);
      if (band > 0.0f) {
        const float motion = fmaxf(fmaxf(slip_v, speed * 0.45f), film);
        f = add_foam(f, breaking_foam(motion, hw) * band, dt);
      }
      const bool on_patch = fabsf(lat) <= p.width * 0.55f + cell * 0.35f && along >= -radius * 0.55f &&
                            along <= radius * 0.35f;
      if (on_patch && hw > 0.02f) {
        const float shear_v = fmaxf(slip_v, film * 0.35f);
        // lift_silt (:6691-6702)
        if (fin(shear_v) && shear_v > 0.0f) {
          const float tau = 1000.0f * 0.035f * shear_v * shear_v;
          if (tau > 6.0f) {
            const float rate = fminf(0.008f * (tau / 6.0f - 1.0f), 0.015f);
            lifted += (double)(add_silt(&tv, fminf(rate * dt, 0.0025f)) * area);
          }
        }
      }
    }
    if (shore[idx]) {
      float prod, rate;
      shore_cell(water, uw, vw, idx, &prod, &rate);
      f = add_foam(f, prod, dt);
      if (rate > 0.0f) lifted += (double)(add_silt(&tv, fminf(rate * dt, 0.0012f)) * area);
    }
    const float h = water[idx] + slurry[idx];
    if (fin(h) && !(h < 0.004f)) {
      const float u = uw[idx], v = vw[idx];
      if (fin(u) && fin(v) && hypot_ref(u, v) > 0.05f) f = add_foam(f, breaking_foam(hypot_ref(u, v), h), dt);
    }
    foam[idx] = f;
    t[idx] = tv;
  }
  wave_add(stirred, lifted);
}

// spawn_breaking_crest (:6563-6600)
__global__ void k_tr_crest(float* __restrict__ foam, const float* __restrict__ water, const float* __restrict__ slurry,
                           const uint8_t* __restrict__ wet, const TracerState* s, float dt) {
  const int idx = cell_index();
  if (idx >= R * R || !s->run || s->nwet == 0 || !wet[idx]) return;
  const int i = idx % R, j = idx / R;
  const float h = water[idx] + slurry[idx];
  if (!fin(h) || h < 0.004f) return;
  float sum = 0.0f, count = 0.0f;
  const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
  for (int q = 0; q < 4; q++) {
    const int ni = i + di[q], nj = j + dj[q];
    if (ni < 0 || nj < 0 || ni >= R || nj >= R) continue;
    const float nh = water[nj * R + ni] + slurry[nj * R + ni];
    if (fin(nh)) {
      sum = sum + fmaxf(nh, 0.0f);
      count = count + 1.0f;
    }
  }
  if (count <= 0.0f) return;
  const float excess = h - sum / count;
  if (excess <= 0.0f) return;
  foam[idx] = add_foam(foam[idx], smooth_step(0.004f, 0.018f, excess), dt);
}

// clear_dry_tracers (:6717-6749): foam half (wet cells), then the silt of
// dry cells in the span lands.
__global__ void k_tr_clear_foam(float* __restrict__ foam, const float* __restrict__ water,
                                const float* __restrict__ slurry, const uint8_t* __restrict__ wet,
                                const TracerState* s, int note) {
  const int idx = cell_index();
  const bool go = note ? (s->run && s->nwet > 0) : proceed(s);
  if (idx >= R * R || !go || !wet[idx]) return;
  const float h = water[idx] + slurry[idx];
  if (!fin(h) || h < 0.004f || !fin(foam[idx])) foam[idx] = 0.0f;
}

__global__ void k_tr_clear_silt(float* __restrict__ t, const float* __restrict__ water,
                                const float* __restrict__ slurry, const TracerState* s, int note, double* settled) {
  const int idx = cell_index();
  double landed = 0.0;
  int i0, i1, j0, j1;
  const bool go = note ? (s->run && s->nwet > 0) : proceed(s);
  if (idx < R * R && go && grown(s->silt_now, &i0, &i1, &j0, &j1)) {
    const int i = idx % R, j = idx / R;
    if (i >= i0 && i <= i1 && j >= j0 && j <= j1) {
      const float h = water[idx] + slurry[idx];
      const float tv = t[idx];
      if (!fin(h) || h < 1.0e-5f) {
        if (tv > 0.0f) {
          landed = (double)(tv * (kCoarseCell * kCoarseCell));
          t[idx] = 0.0f;
        }
      } else if (!fin(tv)) {
        t[idx] = 0.0f;
      }
    }
  }
  wave_add(settled, landed);
}

// sync_foam_live (:6339-6350)
__global__ void k_tr_live(const float* __restrict__ foam, const uint8_t* __restrict__ wet, uint8_t* __restrict__ live,
                          const TracerState* s, int note) {
  const int idx = cell_index();
  const bool go = note ? (s->run && s->nwet > 0) : proceed(s);
  if (idx >= R * R || !go) return;
  const float f = foam[idx];
  live[idx] = wet[idx] && fin(f) && f > 1.0e-4f;
}

constexpr int kBlk = 256;
inline dim3 cells_grid() { return dim3((unsigned)((R * R + kBlk - 1) / kBlk)); }

}  // namespace

void tracers_scan(const TracerPlanes& tr, TracerState* s, hipStream_t st) {
  hipLaunchKernelGGL(k_tr_reset, dim3(1), dim3(64), 0, st, s);
  hipLaunchKernelGGL(k_tr_scan, cells_grid(), dim3(kBlk), 0, st, tr.turbid, tr.foam_live, s);
}

void silt_span_now(const TracerPlanes& tr, TracerState* s, hipStream_t st) {
  hipLaunchKernelGGL(k_tr_span_reset, dim3(1), dim3(64), 0, st, &s->silt_now);
  hipLaunchKernelGGL(k_tr_span, cells_grid(), dim3(kBlk), 0, st, tr.turbid, &s->silt_now);
}

void tracer_step(const TracerPlanes& tr, const CorePlanes& c, TracerState* s, const TracerPushes& tp, int wake,
                 const SpanBox& push_box, float dt, float foam_decay, float silt_keep, float ox, float oz,
                 hipStream_t st) {
  const dim3 g = cells_grid(), b(kBlk);
  tracers_scan(tr, s, st);
  hipLaunchKernelGGL(k_tr_box, dim3(1), dim3(64), 0, st, s, wake, push_box, 0);
  hipLaunchKernelGGL(k_tr_collect, g, b, 0, st, c.water, c.slurry, s, tr.wet, tr.shore);
  hipLaunchKernelGGL(k_foam_sample, g, b, 0, st, tr.foam, c.uw, c.vw, tr.wet, s, dt, tr.scratch);
  hipLaunchKernelGGL(k_foam_write, g, b, 0, st, tr.foam, tr.wet, tr.foam_live, s, foam_decay, tr.scratch);
  hipLaunchKernelGGL(k_silt_advect, g, b, 0, st, tr.turbid, c.uw, c.vw, s, dt, tr.scratch + R * R);
  hipLaunchKernelGGL(k_silt_commit, g, b, 0, st, tr.turbid, s, tr.scratch + R * R);
  silt_span_now(tr, s, st);
  hipLaunchKernelGGL(k_silt_settle, g, b, 0, st, tr.turbid, c.water, c.slurry, s, silt_keep, tr.acc + 1);
  hipLaunchKernelGGL(k_tr_spawn, g, b, 0, st, tr.foam, tr.turbid, c.water, c.slurry, c.uw, c.vw, tr.wet, tr.shore, s,
                     tp, dt, ox, oz, tr.acc);
  hipLaunchKernelGGL(k_tr_clear_foam, g, b, 0, st, tr.foam, c.water, c.slurry, tr.wet, s, 0);
  silt_span_now(tr, s, st);
  hipLaunchKernelGGL(k_tr_clear_silt, g, b, 0, st, tr.turbid, c.water, c.slurry, s, 0, tr.acc + 1);
  hipLaunchKernelGGL(k_tr_live, g, b, 0, st, tr.foam, tr.wet, tr.foam_live, s, 0);
}

void tracer_note_foam(const TracerPlanes& tr, const CorePlanes& c, TracerState* s, int wake, const SpanBox& push_box,
                      float dt, hipStream_t st) {
  const dim3 g = cells_grid(), b(kBlk);
  tracers_scan(tr, s, st);
  hipLaunchKernelGGL(k_tr_box, dim3(1), dim3(64), 0, st, s, wake, push_box, 1);
  hipLaunchKernelGGL(k_tr_collect, g, b, 0, st, c.water, c.slurry, s, tr.wet, tr.shore);
  hipLaunchKernelGGL(k_tr_crest, g, b, 0, st, tr.foam, c.water, c.slurry, tr.wet, s, dt);
  hipLaunchKernelGGL(k_tr_clear_foam, g, b, 0, st, tr.foam, c.water, c.slurry, tr.wet, s, 1);
  silt_span_now(tr, s, st);
  hipLaunchKernelGGL(k_tr_clear_silt, g, b, 0, st, tr.turbid, c.water, c.slurry, s, 1, tr.acc + 1);
  hipLaunchKernelGGL(k_tr_live, g, b, 0, st, tr.foam, tr.wet, tr.foam_live, s, 1);
}

}  // namespace silt
