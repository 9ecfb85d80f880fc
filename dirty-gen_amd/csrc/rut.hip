// Bekker rut coupling: the tire-force rut cut into the ground the sheet
// stands on (src/soil.rs stamp_wheel :2526-2879, its loaded lane tread
// branch used by ThroughDrive::stamp_puddle_tires, src/world.rs:4846-4930),
// with the Bekker sinkage of the contact force (bekker_load :1224-1245,
// sinkage_with :932-946, rut_depth_from_force :1036-1061).
//
// The sinkage (18-step bisection of an 8-sample chord integral) and the two
// exponential rates are scalar per contact and are computed on the host; one
// workgroup then stamps the contacts in order, cells of a footprint in
// parallel, and lane 0 sums the cut volume and lays the spoil on the berm
// band in the reference's raster order (over lists compacted in LDS).
#include <hip/hip_runtime.h>
#include <math.h>

#include <vector>

#include "rut.hpp"

namespace silt {

float bekker_load(float sinkage, float radius, float width, float k_eq, float n) {
  if (sinkage <= 1.0e-5f || k_eq <= 0.0f || width <= 0.0f) return 0.0f;
  float z = fminf(sinkage, radius * 0.95f);
  float half = sqrtf(fmaxf(2.0f * radius * z - z * z, 0.0f));
  if (half < 1.0e-5f) return 0.0f;
  const int samples = 8;
  float acc = 0.0f;
  for (int i = 0; i < samples; i++) {
    float t = ((float)i + 0.5f) / (float)samples;
    float x = -half + t * (2.0f * half);
    float y = sqrtf(fmaxf(radius * radius - x * x, 0.0f)) - (radius - z);
    y = fmaxf(y, 0.0f);
    acc += k_eq * powf(y, n);
  }
  float avg_p = acc / (float)samples;
  return avg_p * (2.0f * half) * width;
}

float sinkage_with(float load, float radius, float width, float k_eq, float n) {
  float lo = 0.0f, hi = radius * 0.9f;
  for (int k = 0; k < 18; k++) {
    float mid = 0.5f * (lo + hi);
    float w = bekker_load(mid, radius, width, k_eq, n);
    if (w < load)
      lo = mid;
    else
      hi = mid;
  }
  return 0.5f * (lo + hi);
}

float rut_depth_from_force(float normal, float push, float radius, float width, float k_eq, float n) {
  float nn = fmaxf(normal, 0.0f), p = fmaxf(push, 0.0f);
  float load = sqrtf(nn * nn + p * p);
  if (load < 1.0f || !isfinite(load)) return 0.0f;
  float r = fmaxf(radius, 0.05f);
  float depth = sinkage_with(load, r, fmaxf(width, 0.02f), k_eq, n);
  return fminf(depth, r * 0.85f);
}

float contact_length(float radius, float sinkage) {
  float z = clampf(sinkage, 0.0f, radius * 0.98f);
  return 2.0f * sqrtf(fmaxf(2.0f * radius * z - z * z, 0.0f));
}

namespace {

constexpr int kMaxFoot = 64 * 64;
constexpr int kStampLanes = 256;

// Exclusive prefix sum of v over the workgroup's 256 lanes (wavefront
// shuffles, then the four wavefront totals); *total gets the sum.
__device__ inline int block_exclusive_scan(int v, int* wave_tot, int* total) {
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wave_tot[w] = x;
  __syncthreads();
  int base = 0, all = 0;
#pragma unroll
  for (int k = 0; k < kStampLanes / 64; k++) {
    base += k < w ? wave_tot[k] : 0;
    all += wave_tot[k];
  }
  __syncthreads();
  *total = all;
  return base + x - v;
}

// One workgroup applies a group of `per` consecutive contacts in order
// (stamp_wheel per tire, src/soil.rs:2526-2879); groups whose footprints do
// not overlap run in parallel, one workgroup each.  The footprint's cells
// run in parallel; the two sequential FP32 sums of the reference (the cut
// volume, and the spoil rounds of deposit_spoil) run on lane 0 over lists
// compacted in raster order in LDS, so they add in the reference's order.
// With sheet_bed set, the footprint's sheet bed (bed + placed) and rut
// (max(initial - bed, 0)) are refreshed after each contact.
__global__ __launch_bounds__(kStampLanes) void k_stamp(float* bed, float* placed, const float* initial, int res,
                                                       float cell, float ox, float oz, const RutContact* cs, int nc,
                                                       int per, float* sheet_bed, float* rut) {
  __shared__ float drop_s[kMaxFoot];
  __shared__ unsigned char berm_s[kMaxFoot];
  __shared__ float drops_c[kMaxFoot];   // positive drops, raster order
  __shared__ short berm_k[kMaxFoot];    // berm cells, raster order
  __shared__ float berm_p[kMaxFoot];    // their placed depth
  __shared__ int wave_tot[kStampLanes / 64];
  __shared__ int n_berm_s;
  const int t = (int)threadIdx.x;
  const int q0 = (int)blockIdx.x * per, q1 = min(nc, q0 + per);
  for (int q = q0; q < q1; q++) {
    const RutContact c = cs[q];
    if (c.skip) continue;
    const int W = c.ihi - c.ilo, H = c.jhi - c.jlo;
    const int nfoot = W > 0 && H > 0 ? W * H : 0;
    for (int k = t; k < nfoot; k += (int)blockDim.x) {
      drop_s[k] = 0.0f;
      berm_s[k] = 0;
      int i = c.ilo + k % W, j = c.jlo + k / W;
      float x = ox + ((float)i + 0.5f) * cell, z = oz + ((float)j + 0.5f) * cell;
      float relx = x - c.cx, relz = z - c.cz;
      float lx = relx * c.rx + relz * c.rz;
      float lz = relx * c.fx + relz * c.fz;
      float u = clampf(fabsf(lz) / fmaxf(c.along_reach, 0.05f), 0.0f, 1.0f);
      float along = u < 0.78f ? 1.0f : clampf((1.0f - u) / 0.22f, 0.0f, 1.0f);
      if (along <= 0.0f) continue;
      size_t idx = (size_t)j * res + i;
      float column = bed[idx] + placed[idx];
      float penetration = column - (c.hub_y - c.radius);
      if (penetration > c.sinkage + 0.03f && fabsf(lz) > cell * 0.5f) continue;
      float ax = fabsf(lx);
      float across;
      if (ax <= c.tread_core) {
        across = 1.0f;
      } else if (ax >= c.tread_edge) {
        across = 0.0f;
      } else {
        float tt = (ax - c.tread_core) / fmaxf(c.tread_edge - c.tread_core, 1.0e-3f);
        float s = tt * tt * (3.0f - 2.0f * tt);
        across = 1.0f - s;
      }
      if (across > 0.12f) {
        if (!c.loaded) continue;
        float already = fmaxf(initial[idx] - bed[idx], 0.0f);
        float local_want = c.want * across;
        float local_groove = c.forced * across;
        if (already < local_want) {
          float room = fmaxf(fminf(local_want - already, 0.40f - already), 0.0f);
          float e = already < local_groove ? c.e_fast : c.e_slow;
          float drop = room * e * along;
          if (drop > 0.0f) {
            float from_placed = fminf(drop, placed[idx]);
            placed[idx] = placed[idx] - from_placed;
            bed[idx] = bed[idx] - (drop - from_placed);
            drop_s[k] = drop;
          }
        }
      } else {
        float start = c.tread_core + cell * 2.55f;
        float end = start + cell * 1.35f;
        if (ax > start && ax <= end && along > 0.25f && !(fabsf(x) + 0.06f < fabsf(c.cx))) berm_s[k] = 1;
      }
    }
    __syncthreads();
    // Compact the positive drops and the berm cells in raster order: lane t
    // takes a contiguous run of the footprint.
    const int seg = (nfoot + kStampLanes - 1) / kStampLanes;
    const int k0 = min(t * seg, nfoot), k1 = min(k0 + seg, nfoot);
    int cd = 0, cb = 0;
    for (int k = k0; k < k1; k++) {
      cd += drop_s[k] > 0.0f ? 1 : 0;
      cb += berm_s[k];
    }
    int nd = 0, nb = 0;
    int pd = block_exclusive_scan(cd, wave_tot, &nd);
    int pb = block_exclusive_scan(cb, wave_tot, &nb);
    for (int k = k0; k < k1; k++) {
      if (drop_s[k] > 0.0f) drops_c[pd++] = drop_s[k];
      if (berm_s[k]) {
        berm_k[pb] = (short)k;
        berm_p[pb] = placed[(size_t)(c.jlo + k / W) * res + (c.ilo + k % W)];
        pb++;
      }
    }
    __syncthreads();
    if (t == 0) {
      bool spoil = false;
      // Cut volume in raster order, then deposit_spoil (src/soil.rs:2881-2946)
      // over the berm cells in raster order, at most six rounds to the cap.
      const float area = cell * cell;
      float vol = 0.0f;
      for (int m = 0; m < nd; m++) vol += drops_c[m] * area;
      if (nfoot > 0 && vol > 1.0e-8f && nb > 0) {
        spoil = true;
        float left = vol;
        const float cap = 0.08f;
        // "open" is the berm list minus the cells that reached the cap
        // (drop_s is free again: it holds the open flags).
        for (int m = 0; m < nb; m++) drop_s[m] = 1.0f;
        int no = nb;
        for (int round = 0; round < 6; round++) {
          if (left <= 1.0e-8f || no == 0) break;
          float share = left / ((float)no * area);
          int nn = 0;
          for (int m = 0; m < nb; m++) {
            if (drop_s[m] == 0.0f) continue;
            float room = fmaxf(cap - berm_p[m], 0.0f);
            float add = fminf(share, room);
            if (add > 0.0f) {
              berm_p[m] = berm_p[m] + add;
              left -= add * area;
            }
            if (berm_p[m] < cap - 1.0e-4f)
              nn++;
            else
              drop_s[m] = 0.0f;
          }
          no = nn;
        }
        if (left > 1.0e-8f) {
          float add = left / ((float)nb * area);
          for (int m = 0; m < nb; m++) berm_p[m] = berm_p[m] + add;
        }
      }
      n_berm_s = spoil ? nb : 0;
    }
    __syncthreads();
    const int nw = n_berm_s;
    for (int m = t; m < nw; m += (int)blockDim.x) {
      const int k = berm_k[m];
      placed[(size_t)(c.jlo + k / W) * res + (c.ilo + k % W)] = berm_p[m];
    }
    if (sheet_bed) {
      __syncthreads();
      for (int k = t; k < nfoot; k += (int)blockDim.x) {
        const size_t idx = (size_t)(c.jlo + k / W) * res + (c.ilo + k % W);
        const float b = bed[idx];
        sheet_bed[idx] = b + placed[idx];
        rut[idx] = fmaxf(initial[idx] - b, 0.0f);
      }
    }
    __syncthreads();
  }
}

}  // namespace

int prepare_contacts(const silt_tire_contact* in, size_t n, int res, float cell, float ox, float oz, float dt,
                     std::vector<RutContact>* out) {
  out->assign(n, RutContact{});
  return n ? prepare_contacts_into(in, n, res, cell, ox, oz, dt, out->data()) : 0;
}

int prepare_contacts_into(const silt_tire_contact* in, size_t n, int res, float cell, float ox, float oz, float dt,
                          RutContact* out) {
  const float e_fast = 1.0f - expf(-dt * 48.0f), e_slow = 1.0f - expf(-dt * 2.2f);
  for (size_t q = 0; q < n; q++) {
    const silt_tire_contact& s = in[q];
    RutContact c{};
    c.cx = s.x;
    c.cz = s.z;
    c.hub_y = s.hub_y;
    c.fx = s.forward_x;
    c.fz = s.forward_z;
    c.rx = s.right_x;
    c.rz = s.right_z;
    c.radius = s.radius;
    c.sinkage = s.sinkage;
    c.loaded = s.normal > 1.0f || s.push > 1.0f;
    c.forced = c.loaded ? rut_depth_from_force(s.normal, s.push, s.radius, s.width, s.k_eq, s.n) : 0.0f;
    c.skip = s.sinkage < 0.004f && c.forced < 0.004f;
    float len = fmaxf(contact_length(s.radius, fmaxf(s.sinkage, c.forced)), s.radius * 0.35f);
    c.tread_core = fmaxf(s.width * 0.42f, cell * 0.45f);
    c.tread_edge = c.tread_core + cell * 2.75f;
    float berm_band = cell * 2.4f;
    float half_l = fmaxf(len * 0.5f + cell * 0.25f, cell * 1.05f);
    c.along_reach = half_l + cell * 0.35f;
    float lateral_reach = c.tread_edge + 0.0f + berm_band + cell * 0.4f;
    float reach = c.along_reach + lateral_reach;
    long i0 = (long)floorf((s.x - reach - ox) / cell), i1 = (long)floorf((s.x + reach - ox) / cell);
    long j0 = (long)floorf((s.z - reach - oz) / cell), j1 = (long)floorf((s.z + reach - oz) / cell);
    c.ilo = (int)(i0 > 1 ? i0 : 1);
    c.ihi = (int)(i1 < res - 1 ? i1 : res - 1);
    c.jlo = (int)(j0 > 1 ? j0 : 1);
    c.jhi = (int)(j1 < res - 1 ? j1 : res - 1);
    if (c.ihi < c.ilo) c.ihi = c.ilo;
    if (c.jhi < c.jlo) c.jhi = c.jlo;
    if ((long)(c.ihi - c.ilo) * (long)(c.jhi - c.jlo) > kMaxFoot) return SILT_EINVAL;
    c.want = c.loaded ? fminf(c.forced, 0.40f) : 0.0f;
    c.e_fast = e_fast;
    c.e_slow = e_slow;
    out[q] = c;
  }
  return 0;
}

int rut_stamp(float* bed, float* placed, const float* initial, int res, float cell, float ox, float oz,
              const silt_tire_contact* contacts, size_t n, float dt, RutContact* d_scratch, hipStream_t st,
              int per_group, float* sheet_bed, float* rut) {
  std::vector<RutContact> cs;
  int rc = prepare_contacts(contacts, n, res, cell, ox, oz, dt, &cs);
  if (rc) return rc;
  if (cs.empty()) return 0;
  return rut_stamp_prepared(bed, placed, initial, res, cell, ox, oz, cs.data(), (int)cs.size(), d_scratch, st,
                            per_group, sheet_bed, rut);
}

int rut_stamp_prepared(float* bed, float* placed, const float* initial, int res, float cell, float ox, float oz,
                       const RutContact* host, int nc, RutContact* d_scratch, hipStream_t st, int per_group,
                       float* sheet_bed, float* rut) {
  if (nc <= 0) return 0;
  if (hipMemcpyAsync(d_scratch, host, (size_t)nc * sizeof(RutContact), hipMemcpyHostToDevice, st) != hipSuccess)
    return SILT_EDEVICE;
  const int per = per_group > 0 && per_group < nc ? per_group : nc;
  hipLaunchKernelGGL(k_stamp, dim3((unsigned)((nc + per - 1) / per)), dim3(kStampLanes), 0, st, bed, placed, initial,
                     res, cell, ox, oz, d_scratch, nc, per, sheet_bed, rut);
  return hipGetLastError() == hipSuccess ? 0 : SILT_EDEVICE;
}

}  // namespace silt
