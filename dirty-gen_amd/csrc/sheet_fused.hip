// Fused streaming sheet substep for gfx950.
//
// One workgroup of 256 lanes owns a strip of 256 loaded columns (11 halo
// columns left, 10 right, 235 output columns) and walks down a chunk of rows.
// Every lane is one column.  Each iteration loads one new row and pushes every
// stage of the substep (src/fluid_core.rs:1879-2082) one row further:
//
//   S1  eta = bed + h                                   row r
//   S2  x filters of eta (G, S) and lap(eta)            row r-1
//   S3  z filters -> pressure p                         row r-3
//   S4  gravity (or wind blend) -> snapshot ug, vg      row r-4
//   S5  vorticity w of the snapshot                     row r-5
//   S6  confinement -> uc, vc                           row r-6
//   S7  divergence source b (projection)                row r-7
//   S8  x filters of b                                  row r-7
//   S9  z filters -> correction potential q             row r-9
//   S10 correction kick, damping, speed cap -> uf, vf   row r-10
//   S11 open transport along x -> h1                    row r-10
//   S12 open transport along z, dry cut, store          row r-11
//
// Horizontal neighbours go through LDS row rings (one barrier between a
// producer and a consumer); vertical neighbours of a lane's own column are
// register windows or the same rings.  The reference clamps every stencil at
// the grid edge; here the producer of a row writes its edge values into ghost
// columns and ghost rows, so every read is a fixed offset.  Work boxes enter as
// masks: filter planes are zero outside the boxes, forces act on wet box cells,
// transport faces exist between cells of one box, and the dry cut runs inside
// the boxes (the scratch convention of the exact path).
//
// EXACT = true keeps the reference's operation order with no fused
// multiply-add and IEEE division, and reproduces the exact path bit for bit;
// EXACT = false fuses multiply-adds and multiplies by precomputed reciprocals.
#include "sheet_fused.hpp"

#include <math.h>

#include <vector>

namespace silt {

namespace {

constexpr int kLanes = 256;
constexpr int kHaloL = 11;
constexpr int kHaloR = 10;
constexpr int kOutW = kLanes - kHaloL - kHaloR;  // 235
constexpr int kLag = 11;                        // S12 trails S1 by 11 rows
constexpr int kRow = 264;                       // LDS row stride (256 + ghosts)
constexpr int kMaxBoxes = 8;

// Ring depths (rows) and their offsets in the LDS row pool.
constexpr int D_ETA = 5, D_LAPE = 7, D_P = 4, D_UG = 4, D_VG = 4, D_W = 4, D_UC = 5, D_VC = 6, D_B = 1, D_Q = 4,
              D_UF = 2, D_HX = 1, D_GT = 4;
constexpr int O_ETA = 0;
constexpr int O_LAPE = O_ETA + D_ETA;
constexpr int O_P = O_LAPE + D_LAPE;
constexpr int O_UG = O_P + D_P;
constexpr int O_VG = O_UG + D_UG;
constexpr int O_W = O_VG + D_VG;
constexpr int O_UC = O_W + D_W;
constexpr int O_VC = O_UC + D_UC;
constexpr int O_B = O_VC + D_VC;
constexpr int O_Q = O_B + D_B;
constexpr int O_UF = O_Q + D_Q;
constexpr int O_HX = O_UF + D_UF;
constexpr int O_GT = O_HX + D_HX;
constexpr int kRows = O_GT + D_GT;
constexpr int kBias = 420 * 16;  // multiple of every depth, keeps row + bias >= 0

struct BoxList {
  int n;
  int i0[kMaxBoxes], i1[kMaxBoxes], j0[kMaxBoxes], j1[kMaxBoxes];
};

struct LayerArgs {
  const float* h;
  const float* u;
  const float* v;
  const float* nu;
  const uint8_t* gate;
  float* ho;
  float* uo;
  float* vo;
  StepArgs a;
  BoxList boxes;
  SheetStats* stats;    // plan statistics of the output (zeroed before launch), or null
  SheetTotals* totals;  // running FP64 ledger lines, or null
  float inv2c;    // 1 / (2 cell)
  float invc2;    // 1 / (cell cell)
  float dtc;      // dt / cell
  float kdt;      // 0.012 dt
  float cdt;      // vorticity dt
};

struct FusedArgs {
  const float* bed;
  int res;
  int strips;
  int chunk;   // output rows per workgroup
  int chunks;
  int tiles;   // strips * chunks per layer
  LayerArgs layer[2];
};

template <bool EXACT>
struct Ar {
  static __device__ __forceinline__ float fma_(float a, float b, float c) {
    if constexpr (EXACT)
      return a * b + c;
    else
      return __builtin_fmaf(a, b, c);
  }
  // x / d, with r = 1 / d on the fast path.
  static __device__ __forceinline__ float div(float x, float d, float r) {
    if constexpr (EXACT)
      return x / d;
    else
      return x * r;
  }
};

__device__ __forceinline__ int slot(int row, int depth) { return (row + kBias) % depth; }

}  // namespace

template <bool EXACT, bool PROJECT, bool WIND>
__global__ __launch_bounds__(kLanes) void k_sheet_stream(FusedArgs fa) {
  using A = Ar<EXACT>;
  __shared__ float lds[kRows * kRow];

  const int res = fa.res;
  // XCD-aware tile order: the 8 XCDs take workgroups round robin, so when the
  // count divides evenly each XCD gets a contiguous run of tiles (neighbouring
  // strips and chunks share halo rows in that XCD's L2).
  const int total = (int)gridDim.x;
  const int wg = (int)blockIdx.x;
  const int tile_lin = (total % 8 == 0) ? (wg % 8) * (total / 8) + wg / 8 : wg;
  const int per_layer = fa.tiles;
  const int layer = tile_lin / per_layer;
  const int tile = tile_lin - layer * per_layer;
  const LayerArgs& L = fa.layer[layer];
  const StepArgs& a = L.a;
  const int strip = tile % fa.strips;
  const int chunk = tile / fa.strips;
  const int x0 = strip * kOutW;              // first output column
  const int y0 = chunk * fa.chunk;           // first output row
  const int y1 = min(res, y0 + fa.chunk);    // one past the last output row
  const int t = (int)threadIdx.x;
  const int col = x0 - kHaloL + t;           // this lane's grid column
  const bool col_in = col >= 0 && col < res;
  const int colc = col < 0 ? 0 : (col >= res ? res - 1 : col);
  const bool left_edge = col == 0, right_edge = col == res - 1;
  const int li = t + 2;  // this lane's LDS column
  const float cell = a.cell;
  const float eps = a.eps;

  // Box membership: bit k set when box k covers the column.
  auto colbits = [&](int c) -> unsigned {
    unsigned m = 0;
    for (int k = 0; k < L.boxes.n; k++)
      if (c >= L.boxes.i0[k] && c <= L.boxes.i1[k]) m |= 1u << k;
    return m;
  };
  const unsigned colmask = colbits(col), colmask_r = colbits(col + 1);
  const unsigned colmask_l1 = colbits(col - 1), colmask_l2 = colbits(col - 2);
  auto rowmask = [&](int y) -> unsigned {
    unsigned m = 0;
    for (int k = 0; k < L.boxes.n; k++)
      if (y >= L.boxes.j0[k] && y <= L.boxes.j1[k]) m |= 1u << k;
    return m;
  };

  auto ring = [&](int base, int depth, int row) -> float* { return lds + (base + slot(row, depth)) * kRow; };
  // Store one lane value of row y into a ring, with the clamp ghosts.
  auto put = [&](int base, int depth, int y, float val, bool ghost_rows) {
    if (y < 0 || y >= res || !col_in) return;
    float* r = ring(base, depth, y);
    r[li] = val;
    if (left_edge) {
      r[li - 1] = val;
      r[li - 2] = val;
    }
    if (right_edge) {
      r[li + 1] = val;
      r[li + 2] = val;
    }
    if (ghost_rows) {
      if (y == 0) {
        float* g = ring(base, depth, -1);
        g[li] = val;
        if (left_edge) g[li - 1] = g[li - 2] = val;
        if (right_edge) g[li + 1] = g[li + 2] = val;
      }
      if (y == res - 1) {
        float* g = ring(base, depth, res);
        g[li] = val;
        if (left_edge) g[li - 1] = g[li - 2] = val;
        if (right_edge) g[li + 1] = g[li + 2] = val;
      }
    }
  };
  auto rowptr = [&](const float* p, int y) -> const float* {
    int yc = y < 0 ? 0 : (y >= res ? res - 1 : y);
    return p + (size_t)yc * res;
  };

  const float inv2c = L.inv2c;
  const float two = 2.0f * cell;
  constexpr float kG[5] = {0.06f, 0.24f, 0.40f, 0.24f, 0.06f};
  constexpr float kS[5] = {0.02f, 0.08f, 0.80f, 0.08f, 0.02f};
  auto acc5 = [&](float a0, float a1, float a2, float a3, float a4, const float* k) -> float {
    float s = a0 * k[0];
    s = A::fma_(a1, k[1], s);
    s = A::fma_(a2, k[2], s);
    s = A::fma_(a3, k[3], s);
    s = A::fma_(a4, k[4], s);
    return s;
  };
  // z filter at row y from a window w[0..4] = rows y-2..y+2, clamped at the
  // grid's top and bottom rows (accum5_clamped sums from 0.0).  The window is
  // indexed with constants only, so it stays in registers.
  auto zacc = [&](const float* w, int y, const float* k) -> float {
    if (y >= 2 && y + 2 < res) return acc5(w[0], w[1], w[2], w[3], w[4], k);
    auto pick = [&](int d) -> float {
      int yy = y + d;
      yy = yy < 0 ? 0 : (yy >= res ? res - 1 : yy);
      int ix = yy - y + 2;
      return ix <= 0 ? w[0] : ix == 1 ? w[1] : ix == 2 ? w[2] : ix == 3 ? w[3] : w[4];
    };
    float s = 0.0f;
    s = A::fma_(pick(-2), k[0], s);
    s = A::fma_(pick(-1), k[1], s);
    s = A::fma_(pick(0), k[2], s);
    s = A::fma_(pick(1), k[3], s);
    s = A::fma_(pick(2), k[4], s);
    return s;
  };

  // Register windows: x-filtered eta and b, rows y-2..y+2 of their z filter;
  // transported depth h1 and final vf, rows y-2..y+1 of the z transport.
  float gx[5] = {0, 0, 0, 0, 0}, sx[5] = {0, 0, 0, 0, 0};
  float gbx[5] = {0, 0, 0, 0, 0}, sbx[5] = {0, 0, 0, 0, 0};
  float h1w[4] = {0, 0, 0, 0}, vfw[4] = {0, 0, 0, 0};
  unsigned rmw[4] = {0, 0, 0, 0};  // row box masks for rows y-2..y+1

  // Epilogue accumulators over the cells this workgroup stores: FP64 volume,
  // dry-cut volume and transport residue (what FP32 rounding and the max(0)
  // clamp added beyond the fluxes), plus the wet bounding box, wet count and
  // largest squared speed the next step's plan needs.
  const bool want_stats = L.stats != nullptr;
  double acc_vol = 0.0, acc_cut = 0.0, acc_res = 0.0;
  unsigned acc_wet = 0;
  int acc_i0 = 0x7fffffff, acc_i1 = -1, acc_j0 = 0x7fffffff, acc_j1 = -1;
  float acc_s2 = 0.0f;
  const bool out_col = t >= kHaloL && t < kHaloL + kOutW && col_in;

  const int r_begin = y0 - kHaloL;
  const int r_end = y1 - 1 + kLag;
  const bool nu_uni = a.nu_uniform != 0;
  const bool gate_uni = a.gate_uniform != 0;

  for (int r = r_begin; r <= r_end; r++) {
    // ---- S1: eta at row r ------------------------------------------------
    {
      const int y = r;
      float hv = rowptr(L.h, y)[colc];
      float bv = rowptr(fa.bed, y)[colc];
      put(O_ETA, D_ETA, y, bv + hv, true);
    }
    __syncthreads();
    // ---- S2: x filters of eta, lap(eta) at row r-1 -------------------------
    {
      const int y = r - 1;
      const float* e = ring(O_ETA, D_ETA, y);
      float g = 0.0f, s = 0.0f;
      const unsigned rm = rowmask(y);
      if (!WIND) {
        float e0 = e[li - 2], e1 = e[li - 1], e2 = e[li], e3 = e[li + 1], e4 = e[li + 2];
        bool inside = (colmask & rm) != 0;
        // accum5 (interior) and accum5_clamped (edges) differ only in a
        // leading 0.0 +, which leaves the value unchanged.
        g = inside ? acc5(e0, e1, e2, e3, e4, kG) : 0.0f;
        s = inside ? acc5(e0, e1, e2, e3, e4, kS) : 0.0f;
      }
      for (int k = 0; k < 4; k++) {
        gx[k] = gx[k + 1];
        sx[k] = sx[k + 1];
      }
      gx[4] = g;
      sx[4] = s;
      if (PROJECT) {
        const float* en = ring(O_ETA, D_ETA, y + 1);
        const float* ep = ring(O_ETA, D_ETA, y - 1);
        float sum = e[li + 1] + e[li - 1] + en[li] + ep[li] - 4.0f * e[li];
        float lp = A::div(sum, cell * cell, L.invc2);
        put(O_LAPE, D_LAPE, y, lp, false);
      }
    }
    // ---- S3: pressure at row r-3 ------------------------------------------
    if (!WIND) {
      const int y = r - 3;
      float pv = 0.0f;
      if ((colmask & rowmask(y)) != 0) {
        float gz = zacc(gx, y, kG), sz = zacc(sx, y, kS);
        float filtered = A::fma_(0.38f, sz, 0.62f * gz);
        float ev = ring(O_ETA, D_ETA, y)[li];
        pv = A::fma_(0.55f, ev, 0.45f * filtered);
      }
      put(O_P, D_P, y, pv, true);
    }
    __syncthreads();
    // ---- S4: gravity / wind blend at row r-4 -> snapshot ------------------
    {
      const int y = r - 4;
      float hv = rowptr(L.h, y)[colc];
      float uu = rowptr(L.u, y)[colc];
      float vv = rowptr(L.v, y)[colc];
      bool inside = (colmask & rowmask(y)) != 0;
      if (inside && hv > eps) {
        if (WIND) {
          uu = A::fma_(a.wx - uu, a.blend, uu);
          vv = A::fma_(a.wz - vv, a.blend, vv);
        } else {
          const float* pr = ring(O_P, D_P, y);
          float px = A::div(pr[li + 1] - pr[li - 1], two, inv2c);
          float pz = A::div(ring(O_P, D_P, y + 1)[li] - ring(O_P, D_P, y - 1)[li], two, inv2c);
          float mob = nu_uni ? a.mob0 : sheet_mobility(rowptr(L.nu, y)[colc]);
          if (EXACT) {
            uu = uu + -kGravity * px * mob * a.dt;
            vv = vv + -kGravity * pz * mob * a.dt;
          } else {
            float k = -kGravity * mob * a.dt;
            uu = __builtin_fmaf(k, px, uu);
            vv = __builtin_fmaf(k, pz, vv);
          }
        }
      }
      put(O_UG, D_UG, y, uu, true);
      put(O_VG, D_VG, y, vv, true);
    }
    __syncthreads();
    // ---- S5: vorticity of the snapshot at row r-5 -------------------------
    if (a.vort > 0.0f) {
      const int y = r - 5;
      const float* vg = ring(O_VG, D_VG, y);
      float dvdx = A::div(vg[li + 1] - vg[li - 1], two, inv2c);
      float dudz = A::div(ring(O_UG, D_UG, y + 1)[li] - ring(O_UG, D_UG, y - 1)[li], two, inv2c);
      put(O_W, D_W, y, dvdx - dudz, true);
    }
    __syncthreads();
    // ---- S6: confinement at row r-6 ---------------------------------------
    {
      const int y = r - 6;
      float uu = ring(O_UG, D_UG, y)[li];
      float vv = ring(O_VG, D_VG, y)[li];
      if (a.vort > 0.0f) {
        float hv = rowptr(L.h, y)[colc];
        bool inside = (colmask & rowmask(y)) != 0;
        if (inside && hv > eps) {
          const float* wr = ring(O_W, D_W, y);
          float w = wr[li];
          if (isfinite(w) && !(fabsf(w) < 1.0e-4f)) {
            float wxp = fabsf(wr[li + 1]), wxm = fabsf(wr[li - 1]);
            float wzp = fabsf(ring(O_W, D_W, y + 1)[li]), wzm = fabsf(ring(O_W, D_W, y - 1)[li]);
            float nx = A::div(wxp - wxm, two, inv2c);
            float nz = A::div(wzp - wzm, two, inv2c);
            float du, dv;
            if (EXACT) {
              du = a.vort * a.dt * (-nz * w);
              dv = a.vort * a.dt * (nx * w);
            } else {
              du = L.cdt * (-nz * w);
              dv = L.cdt * (nx * w);
            }
            if (isfinite(du) && isfinite(dv) && A::fma_(du, du, dv * dv) <= 0.0625f) {
              uu = uu + du;
              vv = vv + dv;
            }
          }
        }
      }
      put(O_UC, D_UC, y, uu, false);
      put(O_VC, D_VC, y, vv, true);
    }
    __syncthreads();
    // ---- S7: divergence source at row r-7 ---------------------------------
    if (PROJECT) {
      const int y = r - 7;
      float bv = 0.0f;
      if ((colmask & rowmask(y)) != 0) {
        const float* ucr = ring(O_UC, D_UC, y);
        float ux = A::div(ucr[li + 1] - ucr[li - 1], two, inv2c);
        float vz = A::div(ring(O_VC, D_VC, y + 1)[li] - ring(O_VC, D_VC, y - 1)[li], two, inv2c);
        float div = ux + vz;
        float lp = ring(O_LAPE, D_LAPE, y)[li];
        float ov = nu_uni ? a.ov0 : overload_of(rowptr(L.nu, y)[colc]);
        if (EXACT)
          bv = div * ov - 0.015f * ov * lp;
        else
          bv = ov * __builtin_fmaf(-0.015f, lp, div);
      }
      put(O_B, D_B, y, bv, false);
    }
    __syncthreads();
    // ---- S8: x filters of b at row r-7; S9: q at row r-9 -------------------
    if (PROJECT) {
      const int y = r - 7;
      const float* br = ring(O_B, D_B, y);
      bool inside = (colmask & rowmask(y)) != 0;
      float b0 = br[li - 2], b1 = br[li - 1], b2 = br[li], b3 = br[li + 1], b4 = br[li + 2];
      float g = inside ? acc5(b0, b1, b2, b3, b4, kG) : 0.0f;
      float s = inside ? acc5(b0, b1, b2, b3, b4, kS) : 0.0f;
      for (int k = 0; k < 4; k++) {
        gbx[k] = gbx[k + 1];
        sbx[k] = sbx[k + 1];
      }
      gbx[4] = g;
      sbx[4] = s;
      const int yq = r - 9;
      float qv = 0.0f;
      if ((colmask & rowmask(yq)) != 0) {
        float gz = zacc(gbx, yq, kG), sz = zacc(sbx, yq, kS);
        qv = A::fma_(0.38f, sz, 0.62f * gz);
      }
      put(O_Q, D_Q, yq, qv, true);
    }
    __syncthreads();
    // ---- S10: correction, damping, cap at row r-10 -------------------------
    {
      const int y = r - 10;
      float uu = ring(O_UC, D_UC, y)[li];
      float vv = ring(O_VC, D_VC, y)[li];
      float hv = rowptr(L.h, y)[colc];
      bool inside = (colmask & rowmask(y)) != 0;
      if (inside && hv > eps) {
        if (PROJECT) {
          const float* qr = ring(O_Q, D_Q, y);
          float qx = A::div(qr[li + 1] - qr[li - 1], two, inv2c);
          float qz = A::div(ring(O_Q, D_Q, y + 1)[li] - ring(O_Q, D_Q, y - 1)[li], two, inv2c);
          if (EXACT) {
            uu = uu + 0.012f * a.dt * qx;
            vv = vv + 0.012f * a.dt * qz;
          } else {
            uu = __builtin_fmaf(L.kdt, qx, uu);
            vv = __builtin_fmaf(L.kdt, qz, vv);
          }
        }
        float decay = nu_uni ? a.decay0 : expf(-sheet_damp(rowptr(L.nu, y)[colc]) * a.dt);
        uu = finite_speed(uu * decay, a.lim);
        vv = finite_speed(vv * decay, a.lim);
      }
      put(O_UF, D_UF, y, uu, false);
      // vf stays in this lane: the z transport reads only its own column.
      for (int k = 0; k < 3; k++) vfw[k] = vfw[k + 1];
      vfw[3] = vv;
      put(O_HX, D_HX, y, hv, false);
      if (!gate_uni) {
        int yc = y < 0 ? 0 : (y >= res ? res - 1 : y);
        put(O_GT, D_GT, y, (float)L.gate[(size_t)yc * res + colc], false);
      }
    }
    __syncthreads();
    // ---- S11: transport along x at row r-10 --------------------------------
    const int yx = r - 10;
    const unsigned rmx = rowmask(yx);
    {
      const float* hr = ring(O_HX, D_HX, yx);
      const float* ur = ring(O_UF, D_UF, yx);
      const float* gr = ring(O_GT, D_GT, yx);
      auto gate_at = [&](int d) -> bool { return gate_uni ? true : gr[li + d] != 0.0f; };
      // Edge (c, c+1) exists when both cells lie in one box.
      const bool e_m2 = (colmask_l2 & colmask_l1 & rmx) != 0;  // edge (col-2, col-1)
      const bool e_m1 = (colmask_l1 & colmask & rmx) != 0;     // edge (col-1, col)
      const bool e_0 = (colmask & colmask_r & rmx) != 0;  // edge (col, col+1)
      const float dtc = L.dtc;
      // Left-going flux of edge (k, k+1) (donor k+1), k = col + d.
      auto fl = [&](int d) -> float {
        float speed = 0.5f * (ur[li + d] + ur[li + d + 1]);
        float hd = hr[li + d + 1], hrv = hr[li + d];
        bool live = !(fabsf(speed) < 1.0e-5f);
        bool gate_shut = !gate_at(d) && hrv <= 1.0e-6f;
        float frac = fminf(EXACT ? fabsf(speed) * a.dt / cell : fabsf(speed) * dtc, 0.45f);
        float dh = fminf(hd * frac, hd);
        bool ok = live && !(hd <= 1.0e-8f) && !gate_shut && !(dh <= 1.0e-8f) && !(speed > 0.0f);
        return ok ? dh : 0.0f;
      };
      // Right-going flux of edge (k, k+1) (donor k) with the donor's removed
      // sum holding the left-going flux of edge (k-1, k).
      auto fr = [&](int d, float prev) -> float {
        float speed = 0.5f * (ur[li + d] + ur[li + d + 1]);
        float hd = hr[li + d], hrv = hr[li + d + 1];
        bool live = !(fabsf(speed) < 1.0e-5f);
        float room = hd - (0.0f + prev);
        bool gate_shut = !gate_at(d + 1) && hrv <= 1.0e-6f;
        float frac = fminf(EXACT ? fabsf(speed) * a.dt / cell : fabsf(speed) * dtc, 0.45f);
        float dh = fminf(hd * frac, room);
        bool ok = live && !(room <= 1.0e-8f) && !gate_shut && !(dh <= 1.0e-8f) && speed > 0.0f;
        return ok ? dh : 0.0f;
      };
      float fl_m2 = e_m2 ? fl(-2) : 0.0f;
      float fl_m1 = e_m1 ? fl(-1) : 0.0f;
      float fr_m1 = e_m1 ? fr(-1, fl_m2) : 0.0f;
      float fl_0 = e_0 ? fl(0) : 0.0f;
      float fr_0 = e_0 ? fr(0, fl_m1) : 0.0f;
      float rem = (0.0f + fl_m1) + fr_0;
      float add = (0.0f + fr_m1) + fl_0;
      float hv = hr[li];
      float moved = fmaxf(hv - rem + add, 0.0f);
      float h1 = (rem > 0.0f || add > 0.0f) ? moved : hv;
      if (want_stats && out_col && yx >= y0 && yx < y1)
        acc_res += ((double)h1 - (double)hv) - ((double)add - (double)rem);
      for (int k = 0; k < 3; k++) {
        h1w[k] = h1w[k + 1];
        rmw[k] = rmw[k + 1];
      }
      h1w[3] = h1;
      rmw[3] = rmx;
    }
    // ---- S12: transport along z, dry cut, store at row r-11 ----------------
    {
      const int y = r - 11;
      // Windows hold rows y-2..y+1 at indices 0..3.
      const float* g_m1 = ring(O_GT, D_GT, y - 1);
      const float* g_0 = ring(O_GT, D_GT, y);
      const float* g_p1 = ring(O_GT, D_GT, y + 1);
      const bool e_m2 = (colmask & rmw[0] & rmw[1]) != 0;  // edge (y-2, y-1)
      const bool e_m1 = (colmask & rmw[1] & rmw[2]) != 0;  // edge (y-1, y)
      const bool e_0 = (colmask & rmw[2] & rmw[3]) != 0;   // edge (y, y+1)
      const float dtc = L.dtc;
      auto gate_row = [&](const float* g) -> bool { return gate_uni ? true : g[li] != 0.0f; };
      // Up-going flux of edge (k, k+1) (donor k+1 = window index k+1).
      auto fu = [&](int k, bool recv_gate) -> float {
        float speed = 0.5f * (vfw[k] + vfw[k + 1]);
        float hd = h1w[k + 1], hrv = h1w[k];
        bool live = !(fabsf(speed) < 1.0e-5f);
        bool gate_shut = !recv_gate && hrv <= 1.0e-6f;
        float frac = fminf(EXACT ? fabsf(speed) * a.dt / cell : fabsf(speed) * dtc, 0.45f);
        float dh = fminf(hd * frac, hd);
        bool ok = live && !(hd <= 1.0e-8f) && !gate_shut && !(dh <= 1.0e-8f) && !(speed > 0.0f);
        return ok ? dh : 0.0f;
      };
      auto fd = [&](int k, float prev, bool recv_gate) -> float {
        float speed = 0.5f * (vfw[k] + vfw[k + 1]);
        float hd = h1w[k], hrv = h1w[k + 1];
        bool live = !(fabsf(speed) < 1.0e-5f);
        float room = hd - (0.0f + prev);
        bool gate_shut = !recv_gate && hrv <= 1.0e-6f;
        float frac = fminf(EXACT ? fabsf(speed) * a.dt / cell : fabsf(speed) * dtc, 0.45f);
        float dh = fminf(hd * frac, room);
        bool ok = live && !(room <= 1.0e-8f) && !gate_shut && !(dh <= 1.0e-8f) && speed > 0.0f;
        return ok ? dh : 0.0f;
      };
      float fu_m2 = e_m2 ? fu(0, gate_row(ring(O_GT, D_GT, y - 2))) : 0.0f;
      float fu_m1 = e_m1 ? fu(1, gate_row(g_m1)) : 0.0f;
      float fd_m1 = e_m1 ? fd(1, fu_m2, gate_row(g_0)) : 0.0f;
      float fu_0 = e_0 ? fu(2, gate_row(g_0)) : 0.0f;
      float fd_0 = e_0 ? fd(2, fu_m1, gate_row(g_p1)) : 0.0f;
      float rem = (0.0f + fu_m1) + fd_0;
      float add = (0.0f + fd_m1) + fu_0;
      float hv = h1w[2];
      float moved = fmaxf(hv - rem + add, 0.0f);
      float h2 = (rem > 0.0f || add > 0.0f) ? moved : hv;
      float uo = ring(O_UF, D_UF, y)[li];
      float vo = vfw[2];
      bool inside = (colmask & rmw[2]) != 0;
      const bool store = out_col && y >= y0 && y < y1;
      if (want_stats && store) acc_res += ((double)h2 - (double)hv) - ((double)add - (double)rem);
      if (inside && eps > 0.0f && h2 < eps) {
        if (want_stats && store) acc_cut += (double)h2;
        h2 = 0.0f;
        uo = 0.0f;
        vo = 0.0f;
      }
      if (store) {
        size_t idx = (size_t)y * res + col;
        L.ho[idx] = h2;
        L.uo[idx] = uo;
        L.vo[idx] = vo;
        if (want_stats) {
          acc_vol += (double)fmaxf(h2, 0.0f);
          if (h2 > eps) {
            acc_wet++;
            acc_i0 = min(acc_i0, col);
            acc_i1 = max(acc_i1, col);
            acc_j0 = min(acc_j0, y);
            acc_j1 = max(acc_j1, y);
          }
          acc_s2 = fmaxf(acc_s2, uo * uo + vo * vo);
        }
      }
    }
    // No barrier here: the next iteration's first writes (eta, lap, p) go to
    // ring slots no stage of this iteration still reads.
  }
  if (want_stats) {
    // Wavefront reduction by shuffles, then one atomic per wave.
    for (int off = 32; off > 0; off >>= 1) {
      acc_vol += __shfl_xor(acc_vol, off);
      acc_cut += __shfl_xor(acc_cut, off);
      acc_res += __shfl_xor(acc_res, off);
      acc_wet += __shfl_xor(acc_wet, off);
      acc_i0 = min(acc_i0, __shfl_xor(acc_i0, off));
      acc_i1 = max(acc_i1, __shfl_xor(acc_i1, off));
      acc_j0 = min(acc_j0, __shfl_xor(acc_j0, off));
      acc_j1 = max(acc_j1, __shfl_xor(acc_j1, off));
      acc_s2 = fmaxf(acc_s2, __shfl_xor(acc_s2, off));
    }
    if ((t & 63) == 0) {
      atomicAdd(&L.stats->vol, acc_vol);
      atomicAdd(&L.stats->wet, (unsigned long long)acc_wet);
      if (acc_wet) {
        atomicMin(&L.stats->i0, acc_i0);
        atomicMax(&L.stats->i1, acc_i1);
        atomicMin(&L.stats->j0, acc_j0);
        atomicMax(&L.stats->j1, acc_j1);
      }
      atomicMax(&L.stats->max_s2_bits, __float_as_uint(acc_s2));
      if (L.totals) {
        atomicAdd(&L.totals->cut, acc_cut);
        atomicAdd(&L.totals->residue, acc_res);
      }
    }
  }
}

namespace {

LayerArgs make_layer(const SheetPlanes& s, float* ho, float* uo, float* vo, const StepArgs& a,
                     const std::vector<Box>& boxes) {
  LayerArgs L{};
  L.h = s.h;
  L.u = s.u;
  L.v = s.v;
  L.nu = s.nu;
  L.gate = s.gate;
  L.ho = ho;
  L.uo = uo;
  L.vo = vo;
  L.a = a;
  L.boxes.n = (int)boxes.size();
  for (size_t k = 0; k < boxes.size() && k < (size_t)kMaxBoxes; k++) {
    L.boxes.i0[k] = boxes[k].i0;
    L.boxes.i1[k] = boxes[k].i1;
    L.boxes.j0[k] = boxes[k].j0;
    L.boxes.j1[k] = boxes[k].j1;
  }
  L.stats = nullptr;
  L.totals = nullptr;
  L.inv2c = 1.0f / (2.0f * a.cell);
  L.invc2 = 1.0f / (a.cell * a.cell);
  L.dtc = a.dt / a.cell;
  L.kdt = 0.012f * a.dt;
  L.cdt = a.vort * a.dt;
  return L;
}

bool boxes_fit(const std::vector<Box>& boxes) {
  if (boxes.empty() || boxes.size() > (size_t)kMaxBoxes) return false;
  // Boxes must neither overlap nor touch, so "both cells in some box" means
  // "both in the same box".
  for (size_t i = 0; i < boxes.size(); i++)
    for (size_t j = i + 1; j < boxes.size(); j++) {
      const Box &p = boxes[i], &q = boxes[j];
      bool apart = p.i1 + 1 < q.i0 || q.i1 + 1 < p.i0 || p.j1 + 1 < q.j0 || q.j1 + 1 < p.j0;
      if (!apart) return false;
    }
  return true;
}

int g_chunk_rows = 0;  // 0: pick automatically

}  // namespace

int fused_chunk_rows(int res) {
  if (g_chunk_rows > 0) return g_chunk_rows;
  // Enough workgroups to give every CU several, few enough to keep the
  // 22-row pipeline fill small.
  if (res >= 2048) return 192;
  if (res >= 512) return 64;
  return 32;
}

void fused_set_chunk_rows(int rows) { g_chunk_rows = rows; }

template <bool EXACT>
static void launch(const FusedArgs& fa, int layers, bool project, bool wind, hipStream_t st) {
  dim3 grid((unsigned)(fa.tiles * layers));
  dim3 blk(kLanes);
  if (wind)
    hipLaunchKernelGGL((k_sheet_stream<EXACT, false, true>), grid, blk, 0, st, fa);
  else if (project)
    hipLaunchKernelGGL((k_sheet_stream<EXACT, true, false>), grid, blk, 0, st, fa);
  else
    hipLaunchKernelGGL((k_sheet_stream<EXACT, false, false>), grid, blk, 0, st, fa);
}

int fused_substep_layers(const FusedLayer* layers, int nlayers, const float* bed, int res, bool exact,
                         hipStream_t st) {
  if (nlayers < 1 || nlayers > 2) return SILT_EINVAL;
  FusedArgs fa{};
  fa.bed = bed;
  fa.res = res;
  fa.strips = (res + kOutW - 1) / kOutW;
  fa.chunk = fused_chunk_rows(res);
  fa.chunks = (res + fa.chunk - 1) / fa.chunk;
  fa.tiles = fa.strips * fa.chunks;
  bool project = layers[0].a.project != 0, wind = layers[0].a.wind != 0;
  for (int k = 0; k < nlayers; k++) {
    const FusedLayer& f = layers[k];
    if ((f.a.project != 0) != project || (f.a.wind != 0) != wind) return SILT_EINVAL;
    if (!boxes_fit(f.boxes)) return SILT_EUNSUPPORTED;
    fa.layer[k] = make_layer(f.in, f.ho, f.uo, f.vo, f.a, f.boxes);
    fa.layer[k].stats = f.stats;
    fa.layer[k].totals = f.totals;
  }
  if (exact)
    launch<true>(fa, nlayers, project, wind, st);
  else
    launch<false>(fa, nlayers, project, wind, st);
  return hipGetLastError() == hipSuccess ? 0 : SILT_EDEVICE;
}

// One planned step of one layer: nsub fused substeps ping-ponging between
// the caller's planes and the workspace copies.  Plans the fused kernel does
// not cover (pace 3 slices, Gauss-Seidel sweeps, overlapping boxes) return
// SILT_EUNSUPPORTED so the caller can take the exact path.
int fused_step(const SheetPlanes& s, const StepPlan& plan, int res, float cell, float dt, const silt_drive& d,
               const silt_sheet_opts& o, bool gate_uniform, bool nu_uniform, float nu0, hipStream_t st) {
  if (plan.pace3 || !boxes_fit(plan.boxes)) return SILT_EUNSUPPORTED;
  float dts = dt / (float)plan.nsub;
  StepArgs a = make_args(res, cell, dts, d, o.eps, o.banks, gate_uniform, nu_uniform, nu0);
  if (a.banks) return SILT_EUNSUPPORTED;
  // Scratch planes reused as the ping-pong buffers.
  float* bh = s.eta;
  float* bu = s.tg;
  float* bv = s.ts;
  size_t bytes = (size_t)res * res * sizeof(float);
  for (int k = 0; k < plan.nsub; k++) {
    bool from_user = (k % 2) == 0;
    FusedLayer f{};
    f.in = s;
    if (!from_user) {
      f.in.h = bh;
      f.in.u = bu;
      f.in.v = bv;
    }
    f.ho = from_user ? bh : s.h;
    f.uo = from_user ? bu : s.u;
    f.vo = from_user ? bv : s.v;
    f.a = a;
    f.boxes = plan.boxes;
    int rc = fused_substep_layers(&f, 1, s.bed, res, o.path == SILT_PATH_FUSED_EXACT, st);
    if (rc != 0) return rc;
  }
  if (plan.nsub % 2 == 1) {
    if (hipMemcpyAsync(s.h, bh, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) return SILT_EDEVICE;
    if (hipMemcpyAsync(s.u, bu, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) return SILT_EDEVICE;
    if (hipMemcpyAsync(s.v, bv, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) return SILT_EDEVICE;
  }
  return plan.nsub;
}

}  // namespace silt
