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
//   S8  x filters of b                                  row r-8
//   S9  z filters -> correction potential q             row r-10
//   S10 correction kick, damping, speed cap -> uf, vf   row r-11
//   S11 open transport along x -> h1                    rows r-12 .. r-14
//   S12 open transport along z, dry cut, store          row r-15
//
// A value a later stage reads across lanes goes through LDS, in a two-slot
// ring (slot = row parity), and is read one iteration after it is written.
// So one barrier per iteration orders every exchange: it separates this
// row's writes from the next row's reads, and this row's reads of a slot from
// the next row's writes to it; every such read of a row is issued at the
// row's start.  The row loop is unrolled by two (six in the interior), so
// the parity and with it every LDS offset is a compile-time constant.  The x
// transport is itself pipelined over three rows: each lane computes the two
// fluxes of its own face (col, col+1) and takes its neighbour's from LDS.
// A value a later stage reads only in its own column (the vertical
// neighbours of every z difference, the lags between stages) stays in
// registers: each lane keeps a short window of the rows it produced.  No
// window is deeper than six rows, so in the six-row interior loop every
// window rotates by register renaming, not by moves.  Global loads run two
// rows ahead of their use.
//
// A workgroup whose window has one viscosity and open gates everywhere (the
// common case) runs a lean copy of the row body without the per-cell law
// branches; a strip inside the grid reads its neighbours' LDS slots
// directly, and a strip that reaches a grid side reads them at clamped
// columns (one base address per offset) instead of storing ghost columns,
// so the edge strips run as fast as the rest.  Three workgroups per CU
// (168 registers per lane) keep the windows out of scratch memory.
//
// The fast path (EXACT = false) also rewrites the arithmetic: the S filter
// is G / 3 + (2/3) delta, so the pressure 0.45 (0.62 G + 0.38 S) + 0.55 eta
// and the correction potential 0.62 G + 0.38 S come from one G filter
// along x and two along z (of the x-filtered rows and of the rows), with
// the separable products expanded; every central difference keeps its
// 1 / (2 cell) in the constant that multiplies it; the vorticity is carried
// as 2 cell w; the finiteness tests that the confinement's magnitude test
// already implies are dropped; the speed cap is a median.
//
// The reference clamps every stencil at the grid edge.  Columns: the lane of
// the first or last grid column also writes its value into two ghost columns.
// Rows: S1 loads clamped rows, so its windows repeat the edge row; windows of
// later stages select the centre value for the missing neighbour of the first
// and last grid row.  Work boxes enter as masks: filter planes are zero
// outside the boxes, forces act on wet box cells, transport faces exist
// between cells of one box, and the dry cut runs inside the boxes (the
// scratch convention of the exact path).
//
// EXACT = true keeps the reference's operation order with no fused
// multiply-add and IEEE division, and reproduces the exact path bit for bit;
// EXACT = false fuses multiply-adds and multiplies by precomputed reciprocals.
// FULL = true is the common plan of one box covering the whole grid: the box
// masks reduce to the grid's own edges.
#include "sheet_fused.hpp"

#include <hip/hip_ext.h>

#include <math.h>

#include <algorithm>
#include <cmath>

#include <map>
#include <string.h>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

namespace silt {

namespace {

constexpr int kLanes = 256;
constexpr int kHaloL = 11;
constexpr int kHaloR = 10;
constexpr int kOutW = kLanes - kHaloL - kHaloR;  // 235
constexpr int kLag = 15;                        // S12 trails S1 by 15 rows
constexpr int kRow = 264;                       // LDS row stride (256 + ghosts)
constexpr int kMaxBoxes = 8;

// LDS rows: a two-slot ring for each value read across lanes (the x fluxes
// fa, fb included), and an eight-slot ring of lap(eta) that each lane reads
// back in its own column six rows later.
constexpr int R_ETA = 0, R_P = 2, R_VG = 4, R_W = 6, R_UC = 8, R_B = 10, R_Q = 12, R_UF = 14, R_HX = 16,
              R_GT = 18, R_FA = 20, R_FB = 22, R_LW = 24;
constexpr int kRows = 32;

// Ring slot of row r - k when r has parity P.
template <int P, int K>
constexpr int kSlot = ((P - K) % 2 + 2) % 2;

struct BoxList {
  int n;
  int i0[kMaxBoxes], i1[kMaxBoxes], j0[kMaxBoxes], j1[kMaxBoxes];
};

// One window of a launch (a layer of a sheet, or one sheet of a batch):
// its scalar fields, which the kernel copies into registers, and its boxes,
// which it reads from memory.
struct LayerHot {
  const float* bed;
  int res;
  int strips;
  int chunk;  // output rows per workgroup
  int tile0;  // first tile of this window in the launch
  const float* h;
  const float* u;
  const float* v;
  const float* nu;
  const uint8_t* gate;
  float* ho;
  float* uo;
  float* vo;
  StepArgs a;
  SheetStats* stats;    // plan statistics of the output (zeroed before launch), or null
  SheetTotals* totals;  // running FP64 ledger lines, or null
  float inv2c;    // 1 / (2 cell)
  float invc2;    // 1 / (cell cell)
  float dtc;      // dt / cell
  float kdt;      // 0.012 dt
  float cdt;      // vorticity dt
};

struct LayerArgs {
  LayerHot hot;
  BoxList boxes;
};

// The descriptors are read through the constant address space: they do not
// change during a launch, so the compiler may load a field wherever it is
// needed instead of holding every field in registers for the whole loop.
using ConstLayer = const __attribute__((address_space(4))) LayerArgs;
using ConstInt = const __attribute__((address_space(4))) int;

// Kernel argument: the windows' descriptors in device memory and, for more
// than two windows, the window of every tile.
struct FusedArgs {
  ConstLayer* table;
  ConstInt* tile_window;
  // Output rows [y0, y1) of every tile (two ints per tile), or null for
  // chunks of the window's uniform height.
  ConstInt* rows;
  // Per-workgroup record for the XCD balance (start and end of the real-time
  // clock, rows the pipeline ran | XCC_ID << 32, tile | HW_ID << 32), or null.
  unsigned long long* clocks;
  int nwin;
};

// RN64(1 / d) of a float divisor d, a type of its own so that a divisor
// cannot be passed where its reciprocal is meant.
struct Recip {
  double v;
};

template <bool EXACT>
struct Ar {
  static __device__ __forceinline__ float fma_(float a, float b, float c) {
    if constexpr (EXACT)
      return a * b + c;
    else
      return __builtin_fmaf(a, b, c);
  }
  // x / d: the IEEE quotient from rd = RN64(1 / d) on the exact path
  // (div_by), x * r with r = RN32(1 / d) on the fast path.
  static __device__ __forceinline__ float div(float x, Recip rd, float r) {
    if constexpr (EXACT)
      return div_by(x, rd.v);
    else
      return x * r;
  }
};

// Squared speed for the plan statistics (one expression for every store).
__device__ __forceinline__ float speed2(float u, float v) { return u * u + v * v; }

// Shift a register window one row on: w[0] drops, w[N-1] = v.
template <int N>
__device__ __forceinline__ void push(float (&w)[N], float v) {
#pragma unroll
  for (int k = 0; k < N - 1; k++) w[k] = w[k + 1];
  w[N - 1] = v;
}

}  // namespace

template <bool EXACT, bool PROJECT, bool WIND, bool FULL>
__global__ __launch_bounds__(kLanes, 3) void k_sheet_stream(FusedArgs fa) {
  using A = Ar<EXACT>;
  __shared__ float lds[kRows * kRow];

  // XCD-aware tile order: the 8 XCDs take workgroups round robin, so XCD x
  // holds workgroups x, x + 8, ...; each XCD gets a contiguous run of tiles
  // (neighbouring strips and chunks share halo rows in that XCD's L2).  With
  // total = 8 q + rem, XCDs below rem hold q + 1 tiles and the rest q.
  const int total = (int)gridDim.x;
  const int wg = (int)blockIdx.x;
  const int xq = total >> 3, xrem = total & 7, xcd = wg & 7;
  const int tile_lin = xcd * xq + min(xcd, xrem) + (wg >> 3);
  const int win =
      fa.nwin <= 2 ? (fa.nwin == 2 && tile_lin >= fa.table[1].hot.tile0 ? 1 : 0) : fa.tile_window[tile_lin];
  ConstLayer& LA = fa.table[win];
  const auto& L = LA.hot;
  const auto& boxes = LA.boxes;
  const auto& a = L.a;
  const int res = L.res;
  const int tile = tile_lin - L.tile0;
  const int strip = tile % L.strips;
  const int chunk = tile / L.strips;
  const int x0 = strip * kOutW;              // first output column
  // First output row and one past the last: uniform chunks, or the rows a
  // table gives each tile (chunk heights sized to each XCD's speed).
  const int y0 = fa.rows ? fa.rows[2 * tile_lin] : chunk * L.chunk;
  const int y1 = fa.rows ? fa.rows[2 * tile_lin + 1] : min(res, y0 + L.chunk);
  const unsigned long long clock0 = fa.clocks ? __builtin_amdgcn_s_memrealtime() : 0ull;
  const int t = (int)threadIdx.x;
  const int col = x0 - kHaloL + t;           // this lane's grid column
  const bool col_in = col >= 0 && col < res;
  const int colc = col < 0 ? 0 : (col >= res ? res - 1 : col);
  const bool left_edge = col == 0, right_edge = col == res - 1;
  // Strips that reach past a grid side carry ghost columns and idle lanes.
  const bool edge_strip = x0 - kHaloL < 0 || x0 - kHaloL + kLanes > res;
  float* const my = lds + (t + 2);  // this lane's LDS column
  const float cell = a.cell;
  const float eps = a.eps;

  // Box membership: bit k set when box k covers the column / row.
  auto colbits = [&](int c) -> unsigned {
    unsigned m = 0;
    for (int k = 0; k < boxes.n; k++)
      if (c >= boxes.i0[k] && c <= boxes.i1[k]) m |= 1u << k;
    return m;
  };
  // With one box over the whole grid every lane counts as inside (lanes off
  // the grid are never stored and never written to LDS) and only the grid's
  // sides end the x faces.
  const unsigned colmask = FULL ? ~0u : colbits(col);
  const unsigned colmask_r = FULL ? (col + 1 < res ? ~0u : 0u) : colbits(col + 1);
  const unsigned colmask_l1 = FULL ? (col >= 1 ? ~0u : 0u) : colbits(col - 1);
  auto rowmask = [&](int y) -> unsigned {
    unsigned m = 0;
    for (int k = 0; k < boxes.n; k++)
      if (y >= boxes.j0[k] && y <= boxes.j1[k]) m |= 1u << k;
    return m;
  };

  // Store this lane's value in LDS row `row` (with the edge ghosts).
  auto put = [&](int row, float val) {
    float* p = my + row * kRow;
    if (!edge_strip) {
      *p = val;
    } else if (col_in) {
      *p = val;
      if (left_edge) {
        p[-1] = val;
        p[-2] = val;
      }
      if (right_edge) {
        p[1] = val;
        p[2] = val;
      }
    }
  };
  auto at = [&](int row, int d) -> float { return my[row * kRow + d]; };
  // Global addresses: the row's base (uniform, scalar registers) plus this
  // lane's byte offset (fixed for the launch), so an access needs no vector
  // address arithmetic.
  const unsigned off_ld = (unsigned)colc * 4u;
  const unsigned off_st = (unsigned)(col < 0 ? 0 : col) * 4u;
  auto row_ld = [&](const float* p, int y) -> float {
    const char* rp = reinterpret_cast<const char*>(p + (size_t)y * (size_t)res);
    return *reinterpret_cast<const float*>(rp + off_ld);
  };
  auto row_st = [&](float* p, int y, float val) {
    char* rp = reinterpret_cast<char*>(p + (size_t)y * (size_t)res);
    *reinterpret_cast<float*>(rp + off_st) = val;
  };
  auto load = [&](const float* p, int y) -> float {
    int yc = y < 0 ? 0 : (y >= res ? res - 1 : y);
    return row_ld(p, yc);
  };

  const float inv2c = L.inv2c;
  // Fast-path constants with the 1 / (2 cell) of each central difference
  // folded in: the gravity kick, the correction kick, the confinement
  // (the vorticity is carried as 2 cell w), its 1e-4 floor, and the x / z
  // transport's half-sum speed.
  const float k_grav = -kGravity * a.mob0 * a.dt * inv2c;
  const float k_grav_nu = -kGravity * a.dt * inv2c;
  const float k_corr = L.kdt * inv2c;
  const float k_conf = -L.cdt * inv2c * inv2c * inv2c;
  const float w_floor = 1.0e-4f / inv2c;
  const float half_dtc = 0.5f * L.dtc;
  // Pressure p = 0.45 (0.62 G + 0.38 S) + 0.55 eta, and the correction
  // potential q = 0.62 G + 0.38 S, with S = G/3 + 2/3 in both directions:
  // a1 GzGx + a2 (Gz + Gx) + a3 for p, b1 .. b3 for q.
  constexpr float kP1 = (float)(0.45 * (0.62 + 0.38 / 9.0)), kP2 = (float)(0.45 * 0.76 / 9.0),
                  kP3 = (float)(0.45 * 1.52 / 9.0 + 0.55);
  constexpr float kQ1 = (float)(0.62 + 0.38 / 9.0), kQ2 = (float)(0.76 / 9.0), kQ3 = (float)(1.52 / 9.0);
  // The exact path's reciprocals of 2 cell, cell and cell^2 (each divisor
  // rounded to float first, as the quotients it stands for divide by it).
  const Recip rd2c{EXACT ? 1.0 / (double)(2.0f * cell) : 0.0};
  const Recip rdcc{EXACT ? 1.0 / (double)(cell * cell) : 0.0};
  const double rdc = EXACT ? 1.0 / (double)cell : 0.0;
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
  // Neighbours of row y in a window holding rows y-1, y, y+1 at m, c, p:
  // the reference clamps the stencil at the first and last grid row.
  auto below = [&](int y, float m, float c) -> float { return y == 0 ? c : m; };
  auto above = [&](int y, float c, float p) -> float { return y == res - 1 ? c : p; };

  // Register windows, oldest row first (rows relative to the iteration's r
  // after its stage has run):
  // FF: the fast path.  Its pressure and correction filters use
  // S = G / 3 + (2/3) delta (the S taps are a third of the G taps plus two
  // thirds of the centre), so the separable 0.62 G + 0.38 S blend needs the
  // G filter along x and two G filters along z (of the x-filtered rows and
  // of the rows themselves, taken as 0 outside the work boxes like the
  // filter planes), and no S planes.
  constexpr bool FF = !EXACT;
  // eta r-3 .. r (r-5 .. r on the FF path); E[EO] is row r-3
  constexpr int EO = FF ? 2 : 0;
  float E[EO + 4] = {};
  float GX[5] = {0, 0, 0, 0, 0}, SX[5] = {0, 0, 0, 0, 0};  // x filters of eta  r-5 .. r-1
  float PW[3] = {0, 0, 0};                                 // p        r-5 .. r-3
  float UG[3] = {0, 0, 0}, VG[3] = {0, 0, 0};              // snapshot r-6 .. r-4
  float WW[3] = {0, 0, 0};                                 // w        r-7 .. r-5
  float UC[6] = {0, 0, 0, 0, 0, 0}, VC[6] = {0, 0, 0, 0, 0, 0};  // confined r-11 .. r-6
  float GB[5] = {0, 0, 0, 0, 0}, SB[5] = {0, 0, 0, 0, 0};  // x filters of b    r-12 .. r-8
  float BW[6] = {0, 0, 0, 0, 0, 0};                        // b (FF)   r-12 .. r-7
  float QW[3] = {0, 0, 0};                                 // q        r-12 .. r-10
  float UF[5] = {0, 0, 0, 0, 0};                           // final u  r-15 .. r-11
  float VF[7] = {0, 0, 0, 0, 0, 0, 0};                     // final v  r-17 .. r-11
  float H1[4] = {0, 0, 0, 0};                              // x-moved h r-17 .. r-14
  float GW[7] = {1, 1, 1, 1, 1, 1, 1};                     // gate     r-17 .. r-11
  unsigned RM[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // RM[k] = row mask of r-k
  // x transport carried between its three rows: this lane's left-going flux
  // of face (col, col+1) and its depth (rows r-13, r-14), the parts of the
  // right-going flux known before the neighbour's (row r-13), and the
  // left-going flux of face (col-1, col) and the right-going flux of face
  // (col, col+1) (row r-14).
  float xa1 = 0.0f, xa2 = 0.0f, xh1 = 0.0f, xh2 = 0.0f, xhdf = 0.0f, xflm1 = 0.0f, xb = 0.0f;
  bool xpre = false;
  // The depth before transport of the x pass's row (r-14) and of the z pass's
  // row (r-15), for the residue.
  float x3_h = 0.0f, x4_h = 0.0f;
  // z fluxes of the edge (y, y+1) of the previous S12 row y, which is the
  // edge (y-1, y) of the next: up-going and down-going.
  float c_fu_0 = 0.0f, c_fd_0 = 0.0f;

  // Epilogue accumulators over the cells this workgroup stores: FP64 volume,
  // dry-cut volume and transport residue (what FP32 rounding and the max(0)
  // clamp added beyond the fluxes), plus the wet bounding box, wet count and
  // largest squared speed the next step's plan needs.  A cell's residue is
  // its depth after both transport passes less its depth before them, in
  // FP64: every face's FP32 flux leaves one cell and enters another, so over
  // the grid the exchanged fluxes cancel and the residues sum to exactly the
  // volume the FP32 updates created or removed.  Cut depths (each below
  // 1e-6 m) are summed per column in FP32 and reduced in FP64.
  const bool want_stats = L.stats != nullptr || L.totals != nullptr;
  double acc_vol = 0.0, acc_res = 0.0;
  float acc_cut = 0.0f;
  float acc_s2 = 0.0f;
  // Wet cells: counted per wavefront from ballots, their rows and lanes
  // kept as scalars (rows come in increasing order).
  unsigned acc_wet = 0;
  int acc_j0 = 0x7fffffff, acc_j1 = -1;
  uint64_t acc_lanes = 0;
  const bool out_col = t >= kHaloL && t < kHaloL + kOutW && col_in;

  // A tile that no work box reaches stores its input unchanged (outside the
  // boxes no force acts, no transport face exists and no dry cut runs), so it
  // only copies its cells and adds them to the statistics.
  bool reached = FULL;
  if (!FULL)
    for (int k = 0; k < boxes.n; k++)
      reached = reached || (boxes.i0[k] < x0 + kOutW && boxes.i1[k] >= x0 && boxes.j0[k] < y1 && boxes.j1[k] >= y0);

  const bool nu_uni = a.nu_uniform != 0;
  const bool gate_uni = a.gate_uniform != 0;
  const bool vort = a.vort > 0.0f;

  // Global loads run two rows ahead of their use, in two register sets by
  // row parity: the loads of row r + 2 are issued at the start of row r and
  // consumed when row r + 2 starts (a row takes about a microsecond, less
  // than a loaded HBM access).
  //   pf_h, pf_b   h and bed of S1's row         (r)
  //   pf_h4, pf_u4, pf_v4   h, u, v of S4's row  (r-4)
  //   pf_h11       h of S10's row                (r-11)
  // S6 reads h of row r-6, which S4 loaded two rows earlier (h6).
  float pf_h[2], pf_b[2], pf_h4[2], pf_u4[2], pf_v4[2], pf_h11[2], h6[2] = {0.0f, 0.0f};
  auto prefetch = [&](auto edge, auto slot, auto lean, int r) __attribute__((always_inline)) {
    constexpr bool EDGE = decltype(edge)::value;
    constexpr int Q = decltype(slot)::value;
    auto ld = [&](const float* p, int y) -> float {
      if constexpr (EDGE) y = y < 0 ? 0 : (y >= res ? res - 1 : y);
      return row_ld(p, y);
    };
    pf_h[Q] = ld(L.h, r);
    pf_b[Q] = ld(L.bed, r);
    pf_h4[Q] = ld(L.h, r - 4);
    pf_u4[Q] = ld(L.u, r - 4);
    pf_v4[Q] = ld(L.v, r - 4);
    pf_h11[Q] = ld(L.h, r - 11);
  };

  // One row of the pipeline.  EDGE = false is the body for rows whose
  // stencils and loads all lie inside the grid: no clamps, no edge selects,
  // and with FULL no row masks.
  // Strips that reach a grid side with one viscosity and open gates (lean
  // mode 2) read their neighbours at clamped columns instead of ghost slots:
  // the lane of column c reads column clamp(c + d), d = -2 .. 2, through its
  // own base address for each d.  Lanes off the grid read the edge column;
  // no stored value depends on them.
  auto lane_of = [&](int c) -> int { return (c < 0 ? 0 : (c >= res ? res - 1 : c)) - (x0 - kHaloL); };
  float* const rb_m2 = lds + 2 + lane_of(col - 2);
  float* const rb_m1 = lds + 2 + lane_of(col - 1);
  float* const rb_p1 = lds + 2 + lane_of(col + 1);
  float* const rb_p2 = lds + 2 + lane_of(col + 2);

  auto row = [&](auto par, auto edge, auto lean, int r) __attribute__((always_inline)) {
    constexpr int P = decltype(par)::value;
    constexpr bool EDGE = decltype(edge)::value;
    // LEAN: a strip inside the grid (no ghost columns) with one viscosity and
    // open gates everywhere, decided once per workgroup.
    constexpr int MODE = decltype(lean)::value;  // 0 general, 1 lean, 2 lean at a grid side
    constexpr bool LEAN = MODE != 0;
    constexpr bool CLAMP = MODE == 2;
    auto at = [&](int rw, int d) -> float {
      if constexpr (CLAMP) {
        if (d == -2) return rb_m2[rw * kRow];
        if (d == -1) return rb_m1[rw * kRow];
        if (d == 1) return rb_p1[rw * kRow];
        if (d == 2) return rb_p2[rw * kRow];
      }
      return my[rw * kRow + d];
    };
    const bool nu_u = LEAN || nu_uni;
    const bool gate_u = LEAN || gate_uni;
    auto putv = [&](int rw, float val) {
      if constexpr (LEAN)
        my[rw * kRow] = val;
      else
        put(rw, val);
    };
    if constexpr (!FULL) {
#pragma unroll
      for (int k = 15; k > 1; k--) RM[k] = RM[k - 1];
      RM[1] = rowmask(r - 1);
    }
    // Row mask of row r - k.
    auto rmk = [&](int k) -> unsigned {
      if constexpr (FULL)
        return !EDGE || (unsigned)(r - k) < (unsigned)res ? 1u : 0u;
      else
        return RM[k];
    };
    auto ld = [&](const float* p, int y) -> float {
      if constexpr (EDGE)
        return load(p, y);
      else
        return row_ld(p, y);
    };
    auto abv = [&](int y, float c, float p) -> float {
      if constexpr (EDGE)
        return above(y, c, p);
      else
        return p;
    };
    auto blw = [&](int y, float m, float c) -> float {
      if constexpr (EDGE)
        return below(y, m, c);
      else
        return m;
    };
    auto zac = [&](const float* w, int y, const float* k) -> float {
      if constexpr (EDGE)
        return zacc(w, y, k);
      else
        return acc5(w[0], w[1], w[2], w[3], w[4], k);
    };

    const float h4 = pf_h4[P], u4 = pf_u4[P], v4 = pf_v4[P], h11 = pf_h11[P];
    // Every value this row reads across lanes was written by the previous
    // row, before the barrier that ended it, into the ring slot this row
    // does not write: all of them are read here, at once, so their latency
    // overlaps (a value whose stage is off this launch is read and unused).
    constexpr int RE = R_ETA + kSlot<P, 1>, RP = R_P + kSlot<P, 4>, RV = R_VG + kSlot<P, 5>,
                  RW = R_W + kSlot<P, 6>, RU = R_UC + kSlot<P, 7>, RB = R_B + kSlot<P, 8>, RQ = R_Q + kSlot<P, 11>,
                  RUF = R_UF + kSlot<P, 12>, RHX = R_HX + kSlot<P, 12>, RGT = R_GT + kSlot<P, 12>,
                  RFA = R_FA + kSlot<P, 13>, RFB = R_FB + kSlot<P, 14>;
    const float l_e0 = at(RE, -2), l_e1 = at(RE, -1), l_e2 = at(RE, 0), l_e3 = at(RE, 1), l_e4 = at(RE, 2);
    const float l_pp = at(RP, 1), l_pm = at(RP, -1);
    const float l_vp = at(RV, 1), l_vm = at(RV, -1);
    const float l_wp = at(RW, 1), l_wm = at(RW, -1);
    const float l_up = at(RU, 1), l_um = at(RU, -1);
    const float l_lp = my[(R_LW + ((r - 7) & 7)) * kRow];
    const float l_b0 = at(RB, -2), l_b1 = at(RB, -1), l_b2 = at(RB, 0), l_b3 = at(RB, 1), l_b4 = at(RB, 2);
    const float l_qp = at(RQ, 1), l_qm = at(RQ, -1);
    const float l_uf0 = at(RUF, 0), l_uf1 = at(RUF, 1), l_hx0 = at(RHX, 0), l_hx1 = at(RHX, 1);
    const float l_g0 = gate_u ? 1.0f : at(RGT, 0), l_g1 = gate_u ? 1.0f : at(RGT, 1);
    const float l_fa = at(RFA, -1), l_fb = at(RFB, -1);
    // ---- S1: eta at row r ------------------------------------------------
    {
      const float ev = pf_b[P] + pf_h[P];
      prefetch(edge, par, lean, r + 2);
      putv(R_ETA + kSlot<P, 0>, ev);
      push(E, ev);
    }
    // ---- S2: x filters of eta, lap(eta) at row r-1 -------------------------
    {
      float g = 0.0f, s = 0.0f;
      if (!WIND) {
        float e0 = l_e0, e1 = l_e1, e2 = l_e2, e3 = l_e3, e4 = l_e4;
        bool inside = (colmask & rmk(1)) != 0;
        // accum5 (interior) and accum5_clamped (edges) differ only in a
        // leading 0.0 +, which leaves the value unchanged.
        if constexpr (EXACT) {
          g = inside ? acc5(e0, e1, e2, e3, e4, kG) : 0.0f;
          s = inside ? acc5(e0, e1, e2, e3, e4, kS) : 0.0f;
        } else {
          const float gv = __builtin_fmaf(0.06f, e0 + e4, __builtin_fmaf(0.24f, e1 + e3, 0.40f * e2));
          g = inside ? gv : 0.0f;
          if constexpr (!FF) s = inside ? __builtin_fmaf(1.0f / 3.0f, gv, (2.0f / 3.0f) * e2) : 0.0f;
        }
      }
      push(GX, g);
      if constexpr (!FF) push(SX, s);
      if (PROJECT) {
        // Rows r and r-2 of this column: E[EO + 3], E[EO + 1] (S1 loaded
        // clamped rows).  The fast path keeps the sum unscaled (1 / cell^2
        // goes into S7).
        float sum = l_e3 + l_e1 + E[EO + 3] + E[EO + 1] - 4.0f * E[EO + 2];
        my[(R_LW + ((r - 1) & 7)) * kRow] = EXACT ? A::div(sum, rdcc, L.invc2) : sum;
      }
    }
    // ---- S3: pressure at row r-3 ------------------------------------------
    if (!WIND) {
      const int y = r - 3;
      float pv;
      if constexpr (FF) {
        // E holds rows r-5 .. r; outside the boxes the rows count as 0
        float gze;
        if constexpr (FULL) {
          gze = zac(E, y, kG);
        } else {
          float em[5];
#pragma unroll
          for (int k = 0; k < 5; k++) em[k] = (colmask & rmk(5 - k)) != 0 ? E[k] : 0.0f;
          gze = zac(em, y, kG);
        }
        const float gzg = zac(GX, y, kG);
        pv = __builtin_fmaf(kP1, gzg, __builtin_fmaf(kP2, gze + GX[2], kP3 * E[2]));
        pv = (colmask & rmk(3)) != 0 ? pv : 0.0f;
      } else {
        float gz = zac(GX, y, kG), sz = zac(SX, y, kS);
        float filtered = A::fma_(0.38f, sz, 0.62f * gz);
        pv = (colmask & rmk(3)) != 0 ? A::fma_(0.55f, E[EO], 0.45f * filtered) : 0.0f;
      }
      putv(R_P + kSlot<P, 3>, pv);
      push(PW, pv);
    }
    // ---- S4: gravity / wind blend at row r-4 -> snapshot ------------------
    {
      const int y = r - 4;
      // The update is computed for every lane and kept on wet box cells:
      // selects, not branches.
      const float hv = h4;
      float uu = u4;
      float vv = v4;
      const bool act = (colmask & rmk(4)) != 0 && hv > eps;
      float un, vn;
      if (WIND) {
        un = A::fma_(a.wx - uu, a.blend, uu);
        vn = A::fma_(a.wz - vv, a.blend, vv);
      } else {
        if (EXACT) {
          float px = A::div(l_pp - l_pm, rd2c, inv2c);
          float pz = A::div(abv(y, PW[1], PW[2]) - blw(y, PW[0], PW[1]), rd2c, inv2c);
          float mob = nu_u ? a.mob0 : sheet_mobility(ld(L.nu, y));
          un = uu + -kGravity * px * mob * a.dt;
          vn = vv + -kGravity * pz * mob * a.dt;
        } else {
          const float k = nu_u ? k_grav : k_grav_nu * sheet_mobility(ld(L.nu, y));
          un = __builtin_fmaf(k, l_pp - l_pm, uu);
          vn = __builtin_fmaf(k, abv(y, PW[1], PW[2]) - blw(y, PW[0], PW[1]), vv);
        }
      }
      uu = act ? un : uu;
      vv = act ? vn : vv;
      putv(R_VG + kSlot<P, 4>, vv);
      push(UG, uu);
      push(VG, vv);
    }
    // ---- S5: vorticity of the snapshot at row r-5 -------------------------
    if (vort) {
      const int y = r - 5;
      float w;
      if constexpr (EXACT) {
        float dvdx = A::div(l_vp - l_vm, rd2c, inv2c);
        float dudz = A::div(abv(y, UG[1], UG[2]) - blw(y, UG[0], UG[1]), rd2c, inv2c);
        w = dvdx - dudz;
      } else {
        // 2 cell w
        w = (l_vp - l_vm) - (abv(y, UG[1], UG[2]) - blw(y, UG[0], UG[1]));
      }
      putv(R_W + kSlot<P, 5>, w);
      push(WW, w);
    }
    // ---- S6: confinement at row r-6 ---------------------------------------
    {
      const int y = r - 6;
      float uu = UG[0];
      float vv = VG[0];
      if (vort) {
        const float hv = h6[P];
        const float w = WW[1];
        float wxp = fabsf(l_wp), wxm = fabsf(l_wm);
        float wzp = fabsf(abv(y, WW[1], WW[2])), wzm = fabsf(blw(y, WW[0], WW[1]));
        float du, dv;
        bool ok;
        if (EXACT) {
          float nx = A::div(wxp - wxm, rd2c, inv2c);
          float nz = A::div(wzp - wzm, rd2c, inv2c);
          du = a.vort * a.dt * (-nz * w);
          dv = a.vort * a.dt * (nx * w);
          ok = (colmask & rmk(6)) != 0 && hv > eps && isfinite(w) && !(fabsf(w) < 1.0e-4f) && isfinite(du) &&
               isfinite(dv) && A::fma_(du, du, dv * dv) <= 0.0625f;
        } else {
          // w is 2 cell w and the gradients 4 cell^2 |w|'s: k_conf holds the
          // 1 / (8 cell^3).  A non-finite w, du or dv fails the magnitude test.
          du = k_conf * ((wzp - wzm) * w);
          dv = -k_conf * ((wxp - wxm) * w);
          ok = (colmask & rmk(6)) != 0 && hv > eps && !(fabsf(w) < w_floor) &&
               __builtin_fmaf(du, du, dv * dv) <= 0.0625f;
        }
        uu = ok ? uu + du : uu;
        vv = ok ? vv + dv : vv;
      }
      putv(R_UC + kSlot<P, 6>, uu);
      push(UC, uu);
      push(VC, vv);
      h6[P] = h4;  // S6's row two rows on
    }
    // ---- S7: divergence source at row r-7 ---------------------------------
    if (PROJECT) {
      const int y = r - 7;
      // VC holds rows r-11..r-6: y-1, y, y+1 at 3, 4, 5.
      float lp = l_lp;
      float ov = nu_u ? a.ov0 : overload_of(ld(L.nu, y));
      float bv;
      if (EXACT) {
        float ux = A::div(l_up - l_um, rd2c, inv2c);
        float vz = A::div(abv(y, VC[4], VC[5]) - blw(y, VC[3], VC[4]), rd2c, inv2c);
        float div = ux + vz;
        bv = div * ov - 0.015f * ov * lp;
      } else {
        // 2 cell div and the unscaled lap sum
        const float div2 = (l_up - l_um) + (abv(y, VC[4], VC[5]) - blw(y, VC[3], VC[4]));
        bv = __builtin_fmaf(-0.015f * ov * L.invc2, lp, (ov * inv2c) * div2);
      }
      bv = (colmask & rmk(7)) != 0 ? bv : 0.0f;
      putv(R_B + kSlot<P, 7>, bv);
      if constexpr (FF) push(BW, bv);
    }
    // ---- S8: x filters of b at row r-8; S9: q at row r-10 ------------------
    if (PROJECT) {
      bool inside = (colmask & rmk(8)) != 0;
      float b0 = l_b0, b1 = l_b1, b2 = l_b2, b3 = l_b3, b4 = l_b4;
      const int yq = r - 10;
      float qv;
      if constexpr (FF) {
        const float gv = __builtin_fmaf(0.06f, b0 + b4, __builtin_fmaf(0.24f, b1 + b3, 0.40f * b2));
        push(GB, inside ? gv : 0.0f);
        const float gzg = zac(GB, yq, kG), gzb = zac(BW, yq, kG);
        qv = __builtin_fmaf(kQ1, gzg, __builtin_fmaf(kQ2, gzb + GB[2], kQ3 * BW[2]));
        qv = (colmask & rmk(10)) != 0 ? qv : 0.0f;
      } else {
        push(GB, inside ? acc5(b0, b1, b2, b3, b4, kG) : 0.0f);
        push(SB, inside ? acc5(b0, b1, b2, b3, b4, kS) : 0.0f);
        float gz = zac(GB, yq, kG), sz = zac(SB, yq, kS);
        qv = (colmask & rmk(10)) != 0 ? A::fma_(0.38f, sz, 0.62f * gz) : 0.0f;
      }
      putv(R_Q + kSlot<P, 10>, qv);
      push(QW, qv);
    }
    // ---- S10: correction, damping, cap at row r-11 -------------------------
    {
      const int y = r - 11;
      float uu = UC[0];
      float vv = VC[0];
      const float hv = h11;
      const bool act = (colmask & rmk(11)) != 0 && hv > eps;
      float un = uu, vn = vv;
      if (PROJECT) {
        if (EXACT) {
          float qx = A::div(l_qp - l_qm, rd2c, inv2c);
          float qz = A::div(abv(y, QW[1], QW[2]) - blw(y, QW[0], QW[1]), rd2c, inv2c);
          un = uu + 0.012f * a.dt * qx;
          vn = vv + 0.012f * a.dt * qz;
        } else {
          un = __builtin_fmaf(k_corr, l_qp - l_qm, uu);
          vn = __builtin_fmaf(k_corr, abv(y, QW[1], QW[2]) - blw(y, QW[0], QW[1]), vv);
        }
      }
      float decay = nu_u ? a.decay0 : expf(-sheet_damp(ld(L.nu, y)) * a.dt);
      if constexpr (EXACT) {
        un = finite_speed(un * decay, a.lim);
        vn = finite_speed(vn * decay, a.lim);
      } else {
        // finite_speed as one median and one finiteness select
        const float ud = un * decay, vd = vn * decay;
        un = isfinite(ud) ? __builtin_amdgcn_fmed3f(ud, -a.lim, a.lim) : 0.0f;
        vn = isfinite(vd) ? __builtin_amdgcn_fmed3f(vd, -a.lim, a.lim) : 0.0f;
      }
      uu = act ? un : uu;
      vv = act ? vn : vv;
      putv(R_UF + kSlot<P, 11>, uu);
      push(UF, uu);
      push(VF, vv);  // the z transport reads only this lane's column
      putv(R_HX + kSlot<P, 11>, hv);
      if (!gate_u) {
        int yc = EDGE ? (y < 0 ? 0 : (y >= res ? res - 1 : y)) : y;
        float gv = (float)L.gate[(size_t)yc * res + colc];
        putv(R_GT + kSlot<P, 11>, gv);
        push(GW, gv);
      }
    }
    // ---- S11: transport along x over rows r-12 .. r-14 ---------------------
    {
      // X1, row r-12: the left-going flux of face (col, col+1) (donor col+1)
      // goes to LDS; of its right-going flux (donor col) all but the room,
      // which needs the neighbour's left-going flux.
      const bool e_0 = (colmask & colmask_r & rmk(12)) != 0;
      // The fast path works with the face's summed speed (twice the mean):
      // the halving is exact, so the tests and the fraction are the same.
      const float speed = EXACT ? 0.5f * (l_uf0 + l_uf1) : l_uf0 + l_uf1;
      const float hc = l_hx0, hr = l_hx1;
      const bool live = !(fabsf(speed) < (EXACT ? 1.0e-5f : 2.0f * 1.0e-5f));
      const float frac = fminf(EXACT ? div_by(fabsf(speed) * a.dt, rdc) : fabsf(speed) * half_dtc, 0.45f);
      const bool shut_l = !gate_u && l_g0 == 0.0f && hc <= 1.0e-6f;
      const bool shut_r = !gate_u && l_g1 == 0.0f && hr <= 1.0e-6f;
      const float dl = fminf(hr * frac, hr);
      // (dl <= hr, so dl > 1e-8 implies the donor's own hr > 1e-8 test,
      // NaNs included; the same holds for every flux below.)
      const bool ok_l = e_0 && live && !shut_l && !(dl <= 1.0e-8f) && !(speed > 0.0f);
      const float fa = ok_l ? dl : 0.0f;
      my[(R_FA + kSlot<P, 12>) * kRow] = fa;
      const bool pre = e_0 && live && !shut_r && speed > 0.0f;
      const float hdf = hc * frac;
      // X2, row r-13: the right-going flux of face (col, col+1), whose donor
      // has already lost the left-going flux of face (col-1, col).
      const bool e_m1b = (colmask_l1 & colmask & rmk(13)) != 0;
      const float fl_m1 = e_m1b ? l_fa : 0.0f;
      const float room = xh1 - (0.0f + fl_m1);
      const float dr = fminf(xhdf, room);
      const float fb = (xpre && !(dr <= 1.0e-8f)) ? dr : 0.0f;
      my[(R_FB + kSlot<P, 13>) * kRow] = fb;
      // X3, row r-14: the depth after the x pass.
      const int yx = r - 14;
      const bool e_m1c = (colmask_l1 & colmask & rmk(14)) != 0;
      const float fr_m1 = e_m1c ? l_fb : 0.0f;
      // The reference sums (0.0 + a) + b; a flux is +0 or positive, never -0,
      // so a + b is the same value.
      const float rem = xflm1 + xb;
      const float add = fr_m1 + xa2;
      const float hv = xh2;
      const float moved = fmaxf(hv - rem + add, 0.0f);
      const float h1 = (rem > 0.0f || add > 0.0f) ? moved : hv;
      (void)yx;
      x3_h = hv;
      push(H1, h1);
      xa2 = xa1;
      xa1 = fa;
      xh2 = xh1;
      xh1 = hc;
      xhdf = hdf;
      xpre = pre;
      xflm1 = fl_m1;
      xb = fb;
    }
    // ---- S12: transport along z, dry cut, store at row r-15 ----------------
    {
      const int y = r - 15;
      // H1, VF and GW hold rows y-2..y+1 at indices 0..3; rmk(15), rmk(14) are
      // the masks of rows y, y+1.
      const bool e_0 = (colmask & rmk(15) & rmk(14)) != 0;  // edge (y, y+1)
      auto gate_ok = [&](float g) -> bool { return gate_u ? true : g != 0.0f; };
      // Up-going flux of edge (k, k+1) (donor k+1 = window index k+1).
      auto fu = [&](int k, bool recv_gate) -> float {
        float speed = EXACT ? 0.5f * (VF[k] + VF[k + 1]) : VF[k] + VF[k + 1];
        float hd = H1[k + 1], hrv = H1[k];
        bool live = !(fabsf(speed) < (EXACT ? 1.0e-5f : 2.0f * 1.0e-5f));
        bool gate_shut = !recv_gate && hrv <= 1.0e-6f;
        float frac = fminf(EXACT ? div_by(fabsf(speed) * a.dt, rdc) : fabsf(speed) * half_dtc, 0.45f);
        float dh = fminf(hd * frac, hd);
        bool ok = live && !gate_shut && !(dh <= 1.0e-8f) && !(speed > 0.0f);
        return ok ? dh : 0.0f;
      };
      auto fd = [&](int k, float prev, bool recv_gate) -> float {
        float speed = EXACT ? 0.5f * (VF[k] + VF[k + 1]) : VF[k] + VF[k + 1];
        float hd = H1[k], hrv = H1[k + 1];
        bool live = !(fabsf(speed) < (EXACT ? 1.0e-5f : 2.0f * 1.0e-5f));
        float room = hd - (0.0f + prev);
        bool gate_shut = !recv_gate && hrv <= 1.0e-6f;
        float frac = fminf(EXACT ? div_by(fabsf(speed) * a.dt, rdc) : fabsf(speed) * half_dtc, 0.45f);
        float dh = fminf(hd * frac, room);
        bool ok = live && !gate_shut && !(dh <= 1.0e-8f) && speed > 0.0f;
        return ok ? dh : 0.0f;
      };
      // The fluxes of the two edges above come from the previous row, where
      // they were the edges (y-1, y) and (y, y+1) of that row.
      const float fu_m1 = c_fu_0, fd_m1 = c_fd_0;
      const float fu_c = fu(2, gate_ok(GW[2]));
      const float fd_c = fd(2, fu_m1, gate_ok(GW[3]));
      const float fu_0 = e_0 ? fu_c : 0.0f;
      const float fd_0 = e_0 ? fd_c : 0.0f;
      c_fu_0 = fu_0;
      c_fd_0 = fd_0;
      float rem = fu_m1 + fd_0;
      float add = fd_m1 + fu_0;
      float hv = H1[2];
      const float h_orig = x4_h;
      x4_h = x3_h;
      float moved = fmaxf(hv - rem + add, 0.0f);
      float h2 = (rem > 0.0f || add > 0.0f) ? moved : hv;
      float uo = UF[0];
      float vo = VF[2];
      const bool inside = (colmask & rmk(15)) != 0;
      const bool store = out_col && y >= y0 && y < y1;
      const bool cut = inside && eps > 0.0f && h2 < eps;
      if (want_stats) {
        // Adding +0.0 for the lanes and rows that are not stored keeps the
        // sums unchanged without a branch.
        if (store) acc_res += (double)h2 - (double)h_orig;
        acc_cut += store && cut ? h2 : 0.0f;
      }
      h2 = cut ? 0.0f : h2;
      uo = cut ? 0.0f : uo;
      vo = cut ? 0.0f : vo;
      if (want_stats) {
        if (store) acc_vol += (double)fmaxf(h2, 0.0f);
        const uint64_t wm = __builtin_amdgcn_ballot_w64(store && h2 > eps);
        acc_wet += (unsigned)__builtin_popcountll(wm);
        acc_lanes |= wm;
        if (wm) {
          acc_j0 = min(acc_j0, y);
          acc_j1 = y;
        }
        acc_s2 = store ? fmaxf(acc_s2, speed2(uo, vo)) : acc_s2;
      }
      if (store) {
        row_st(L.ho, y, h2);
        row_st(L.uo, y, uo);
        row_st(L.vo, y, vo);
      }
    }
    __syncthreads();
  };

  // Rows r_begin..r_end in pairs of (even, odd): a leading row of the wrong
  // parity or a trailing extra row only runs the pipeline over rows that are
  // never stored.
  const int r_begin = (y0 - kLag) & ~1;
  const int r_end = y1 - 1 + kLag;
  using Even = std::integral_constant<int, 0>;
  using Odd = std::integral_constant<int, 1>;
  using Edge = std::true_type;
  using Inside = std::false_type;
  if (reached) {
    auto run = [&](auto lean) __attribute__((always_inline)) {
      prefetch(Edge{}, Even{}, lean, r_begin);
      prefetch(Edge{}, Odd{}, lean, r_begin + 1);
      // Rows r-20 .. r+5 inside the grid cover every stage row, stencil
      // neighbour and load of the pair of rows (r, r+1): those pairs take the
      // interior body.  They form one run, so three loops: top edge, interior,
      // bottom edge.
      auto interior = [&](int r) { return r - 20 >= 0 && r + 5 < res; };
      int r = r_begin;
      for (; r <= r_end && !interior(r); r += 2) {
        row(Even{}, Edge{}, lean, r);
        row(Odd{}, Edge{}, lean, r + 1);
      }
      // Six rows per iteration in the interior: every register window is
      // at most six rows deep, so each rotates by renaming instead of moves.
      for (; r + 4 <= r_end && interior(r) && interior(r + 4); r += 6) {
        row(Even{}, Inside{}, lean, r);
        row(Odd{}, Inside{}, lean, r + 1);
        row(Even{}, Inside{}, lean, r + 2);
        row(Odd{}, Inside{}, lean, r + 3);
        row(Even{}, Inside{}, lean, r + 4);
        row(Odd{}, Inside{}, lean, r + 5);
      }
      for (; r <= r_end && interior(r); r += 2) {
        row(Even{}, Inside{}, lean, r);
        row(Odd{}, Inside{}, lean, r + 1);
      }
      for (; r <= r_end; r += 2) {
        row(Even{}, Edge{}, lean, r);
        row(Odd{}, Edge{}, lean, r + 1);
      }
    };
    if (nu_uni && gate_uni && !edge_strip)
      run(std::integral_constant<int, 1>{});
    else if (nu_uni && gate_uni)
      run(std::integral_constant<int, 2>{});
    else
      run(std::integral_constant<int, 0>{});
  } else {
    // Every lane walks the rows (the wet ballots count whole waves); only the
    // output columns load and store.
    for (int y = y0; y < y1; y++) {
      float hv = 0.0f, uv = 0.0f, vv = 0.0f;
      if (out_col) {
        hv = row_ld(L.h, y);
        uv = row_ld(L.u, y);
        vv = row_ld(L.v, y);
        row_st(L.ho, y, hv);
        row_st(L.uo, y, uv);
        row_st(L.vo, y, vv);
      }
      if (want_stats) {
        if (out_col) acc_vol += (double)fmaxf(hv, 0.0f);
        const uint64_t wm = __builtin_amdgcn_ballot_w64(out_col && hv > eps);
        acc_wet += (unsigned)__builtin_popcountll(wm);
        acc_lanes |= wm;
        if (wm) {
          acc_j0 = min(acc_j0, y);
          acc_j1 = y;
        }
        acc_s2 = out_col ? fmaxf(acc_s2, speed2(uv, vv)) : acc_s2;
      }
    }
  }

  if (want_stats) {
    // Wet lanes of this wave: the columns of the lowest and highest set bit.
    const int lane0 = x0 - kHaloL + (t & ~63);
    int acc_i0 = acc_lanes ? lane0 + __builtin_ctzll(acc_lanes) : 0x7fffffff;
    int acc_i1 = acc_lanes ? lane0 + 63 - __builtin_clzll(acc_lanes) : -1;
    double sum_cut = acc_cut, sum_res = acc_res;
    for (int off = 32; off > 0; off >>= 1) {
      acc_vol += __shfl_xor(acc_vol, off);
      sum_cut += __shfl_xor(sum_cut, off);
      sum_res += __shfl_xor(sum_res, off);
      acc_s2 = fmaxf(acc_s2, __shfl_xor(acc_s2, off));
    }
    if ((t & 63) == 0) {
      if (L.stats) {
        atomicAdd(&L.stats->vol, acc_vol);
        atomicAdd(&L.stats->wet, (unsigned long long)acc_wet);
        if (acc_wet) {
          atomicMin(&L.stats->i0, acc_i0);
          atomicMax(&L.stats->i1, acc_i1);
          atomicMin(&L.stats->j0, acc_j0);
          atomicMax(&L.stats->j1, acc_j1);
        }
        atomicMax(&L.stats->max_s2_bits, __float_as_uint(acc_s2));
      }
      if (L.totals) {
        atomicAdd(&L.totals->cut, sum_cut);
        atomicAdd(&L.totals->residue, sum_res);
      }
    }
  }
  if (fa.clocks) {
    __syncthreads();
    if (t == 0) {
      unsigned long long* c = fa.clocks + 4 * wg;
      c[0] = clock0;
      c[1] = __builtin_amdgcn_s_memrealtime();
      // HW_ID (wave, SIMD, CU, shader engine) and XCC_ID hardware registers.
      const unsigned hw_id = (unsigned)__builtin_amdgcn_s_getreg((31 << 11) | 4);
      const unsigned xcc_id = (unsigned)__builtin_amdgcn_s_getreg((3 << 11) | 20);
      c[2] = (unsigned long long)(reached ? r_end - r_begin + 1 : y1 - y0) | ((unsigned long long)xcc_id << 32);
      c[3] = (unsigned long long)tile_lin | ((unsigned long long)hw_id << 32);
    }
  }
}

namespace {

LayerArgs make_layer(const SheetPlanes& s, float* ho, float* uo, float* vo, const StepArgs& a,
                     const std::vector<Box>& boxes) {
  LayerArgs L{};
  L.hot.h = s.h;
  L.hot.u = s.u;
  L.hot.v = s.v;
  L.hot.nu = s.nu;
  L.hot.gate = s.gate;
  L.hot.ho = ho;
  L.hot.uo = uo;
  L.hot.vo = vo;
  L.hot.a = a;
  L.boxes.n = (int)boxes.size();
  for (size_t k = 0; k < boxes.size() && k < (size_t)kMaxBoxes; k++) {
    L.boxes.i0[k] = boxes[k].i0;
    L.boxes.i1[k] = boxes[k].i1;
    L.boxes.j0[k] = boxes[k].j0;
    L.boxes.j1[k] = boxes[k].j1;
  }
  L.hot.stats = nullptr;
  L.hot.totals = nullptr;
  L.hot.inv2c = 1.0f / (2.0f * a.cell);
  L.hot.invc2 = 1.0f / (a.cell * a.cell);
  L.hot.dtc = a.dt / a.cell;
  L.hot.kdt = 0.012f * a.dt;
  L.hot.cdt = a.vort * a.dt;
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
// Rows per workgroup pinned for one side length and layer count (by the
// autotuner, or for every layer count by silt_set_chunk_rows).
std::mutex g_chunk_mu;
std::map<std::pair<int, int>, int> g_chunk_by_res;

}  // namespace

// Output rows per workgroup.  Every tile pays the 22-row pipeline fill, so
// the chunk is the smallest that keeps all tiles of the launch resident at
// once (one round of workgroups: no tail of late tiles).
int fused_chunk_rows(int res, int strips, int layers, int resident) {
  if (g_chunk_rows > 0) return g_chunk_rows;
  {
    std::lock_guard<std::mutex> lock(g_chunk_mu);
    auto it = g_chunk_by_res.find({res, layers});
    if (it != g_chunk_by_res.end()) return it->second;
  }
  int per_column = resident / (strips * layers);
  if (per_column < 1) per_column = 1;
  int chunk = (res + per_column - 1) / per_column;
  return chunk < 8 ? 8 : chunk;
}

void fused_set_chunk_rows(int rows) { g_chunk_rows = rows; }

void fused_set_chunk_rows_for(int res, int rows, int layers) {
  std::lock_guard<std::mutex> lock(g_chunk_mu);
  for (int l = 1; l <= 2; l++) {
    if (layers != 0 && layers != l) continue;
    if (rows > 0)
      g_chunk_by_res[{res, l}] = rows;
    else
      g_chunk_by_res.erase({res, l});
  }
}

int fused_pinned_chunk_rows(int res, int layers) {
  std::lock_guard<std::mutex> lock(g_chunk_mu);
  for (int l = 2; l >= 1; l--) {
    if (layers != 0 && layers != l) continue;
    auto it = g_chunk_by_res.find({res, l});
    if (it != g_chunk_by_res.end()) return it->second;
  }
  return 0;
}

template <bool EXACT, bool PROJECT, bool WIND, bool FULL>
static int resident_tiles() {
  // Workgroups of this instance the device holds at once (cached per device).
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_sheet_stream<EXACT, PROJECT, WIND, FULL>, kLanes,
                                                     0) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || per_cu < 1 || cus < 1)
      return 256;
    cached[dev] = per_cu * cus;
  }
  return cached[dev];
}

// The instance for a launch: its resident tile count, or the launch itself.
namespace {

// Device copies of the launch descriptors, one buffer per stream: launches on
// one stream run in order, so a buffer is rewritten only after the kernel
// that read it has finished.
struct DescBuf {
  LayerArgs* table = nullptr;
  int* tile_window = nullptr;
  int* rows = nullptr;
  size_t cap_win = 0, cap_tiles = 0, cap_rows = 0;
};
std::mutex g_desc_mu;
std::map<hipStream_t, DescBuf> g_desc;

int upload_descs(const std::vector<LayerArgs>& wins, const std::vector<int>& tile_window,
                 const std::vector<int>& rows, hipStream_t st, FusedArgs* fa) {
  std::lock_guard<std::mutex> lock(g_desc_mu);
  DescBuf& b = g_desc[st];
  if (b.cap_win < wins.size()) {
    if (b.table) (void)hipFree(b.table);
    b.table = nullptr;
    b.cap_win = 0;
    if (hipMalloc(&b.table, wins.size() * sizeof(LayerArgs)) != hipSuccess) return SILT_ENOMEM;
    b.cap_win = wins.size();
  }
  if (b.cap_tiles < tile_window.size()) {
    if (b.tile_window) (void)hipFree(b.tile_window);
    b.tile_window = nullptr;
    b.cap_tiles = 0;
    if (hipMalloc(&b.tile_window, tile_window.size() * sizeof(int)) != hipSuccess) return SILT_ENOMEM;
    b.cap_tiles = tile_window.size();
  }
  if (b.cap_rows < rows.size()) {
    if (b.rows) (void)hipFree(b.rows);
    b.rows = nullptr;
    b.cap_rows = 0;
    if (hipMalloc(&b.rows, rows.size() * sizeof(int)) != hipSuccess) return SILT_ENOMEM;
    b.cap_rows = rows.size();
  }
  // Pageable sources: the copies are staged before the calls return.
  if (hipMemcpyAsync(b.table, wins.data(), wins.size() * sizeof(LayerArgs), hipMemcpyHostToDevice, st) != hipSuccess)
    return SILT_EDEVICE;
  if (!tile_window.empty() && hipMemcpyAsync(b.tile_window, tile_window.data(), tile_window.size() * sizeof(int),
                                             hipMemcpyHostToDevice, st) != hipSuccess)
    return SILT_EDEVICE;
  if (!rows.empty() &&
      hipMemcpyAsync(b.rows, rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess)
    return SILT_EDEVICE;
  fa->table = (ConstLayer*)b.table;
  fa->tile_window = tile_window.empty() ? nullptr : (ConstInt*)b.tile_window;
  fa->rows = rows.empty() ? nullptr : (ConstInt*)b.rows;
  fa->nwin = (int)wins.size();
  return 0;
}

// Rows of every tile with each strip's chunks sized by the weight of the XCD
// that runs them: the boundaries are the strip's cumulative weights scaled
// to the window's side and rounded.  Empty when a tile would get no row.
std::vector<int> balanced_rows(const std::vector<LayerArgs>& wins, int tiles, const float* w) {
  std::vector<int> rows((size_t)2 * tiles);
  std::vector<double> cum;
  for (const LayerArgs& L : wins) {
    const int strips = L.hot.strips, chunks = (L.hot.res + L.hot.chunk - 1) / L.hot.chunk;
    cum.assign((size_t)chunks + 1, 0.0);
    for (int s = 0; s < strips; s++) {
      for (int c = 0; c < chunks; c++) {
        const int tl = L.hot.tile0 + c * strips + s;
        cum[(size_t)c + 1] = cum[(size_t)c] + (double)w[fused_tile_xcd(tl, tiles)];
      }
      int prev = 0;
      for (int c = 0; c < chunks; c++) {
        const int tl = L.hot.tile0 + c * strips + s;
        const int y1 = c + 1 == chunks ? L.hot.res
                                       : (int)std::lround((double)L.hot.res * cum[(size_t)c + 1] / cum[(size_t)chunks]);
        if (y1 <= prev) return {};
        rows[(size_t)2 * tl] = prev;
        rows[(size_t)2 * tl + 1] = y1;
        prev = y1;
      }
    }
  }
  return rows;
}

template <typename K>
static void launch(K kernel, dim3 grid, dim3 blk, hipStream_t st, hipEvent_t e0, hipEvent_t e1, const FusedArgs& fa) {
  if (e0 || e1)
    hipExtLaunchKernelGGL(kernel, grid, blk, 0, st, e0, e1, 0, fa);
  else
    hipLaunchKernelGGL(kernel, grid, blk, 0, st, fa);
}

template <bool EXACT, bool FULL>
static int dispatch(const FusedArgs* fa, int tiles, bool project, bool wind, hipStream_t st, hipEvent_t e0,
                    hipEvent_t e1) {
  if (!fa) {
    if (wind) return resident_tiles<EXACT, false, true, FULL>();
    if (project) return resident_tiles<EXACT, true, false, FULL>();
    return resident_tiles<EXACT, false, false, FULL>();
  }
  dim3 grid((unsigned)tiles);
  dim3 blk(kLanes);
  if (wind)
    launch(k_sheet_stream<EXACT, false, true, FULL>, grid, blk, st, e0, e1, *fa);
  else if (project)
    launch(k_sheet_stream<EXACT, true, false, FULL>, grid, blk, st, e0, e1, *fa);
  else
    launch(k_sheet_stream<EXACT, false, false, FULL>, grid, blk, st, e0, e1, *fa);
  return 0;
}

static int dispatch_any(const FusedArgs* fa, int tiles, bool exact, bool full, bool project, bool wind,
                        hipStream_t st, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  if (exact) return full ? dispatch<true, true>(fa, tiles, project, wind, st, e0, e1)
                         : dispatch<true, false>(fa, tiles, project, wind, st, e0, e1);
  return full ? dispatch<false, true>(fa, tiles, project, wind, st, e0, e1)
              : dispatch<false, false>(fa, tiles, project, wind, st, e0, e1);
}

}  // namespace

size_t fused_desc_size() { return sizeof(LayerArgs); }

int fused_tile_xcd(int tile, int tiles) {
  // Inverse of the kernel's order: XCD x runs tiles [x q + min(x, rem), ...).
  const int q = tiles / kXcds, rem = tiles % kXcds;
  for (int x = kXcds - 1; x > 0; x--)
    if (tile >= x * q + std::min(x, rem)) return x;
  return 0;
}

bool fused_boxes_fit(const std::vector<Box>& boxes) { return boxes_fit(boxes); }

int fused_build(const FusedLayer* wins, int n, bool exact, int rows, FusedLaunch* out, const float* xcd_w) {
  if (n < 1 || !out) return SILT_EINVAL;
  bool project = wins[0].a.project != 0, wind = wins[0].a.wind != 0;
  bool full = true;
  for (int k = 0; k < n; k++) {
    const FusedLayer& f = wins[k];
    if ((f.a.project != 0) != project || (f.a.wind != 0) != wind || f.res < 1 || !f.in.bed) return SILT_EINVAL;
    if (!boxes_fit(f.boxes)) return SILT_EUNSUPPORTED;
    const std::vector<Box>& b = f.boxes;
    full = full && b.size() == 1 && b[0].i0 == 0 && b[0].j0 == 0 && b[0].i1 == f.res - 1 && b[0].j1 == f.res - 1;
  }
  int resident = dispatch_any(nullptr, 0, exact, full, project, wind, nullptr);
  // Rows per workgroup: pinned by the caller, else for one or two windows of
  // one size the occupancy-derived height, else the height that gives about
  // two tiles per resident slot over all windows.
  int chunk = rows;
  if (chunk <= 0) {
    bool same = true;
    for (int k = 1; k < n; k++) same = same && wins[k].res == wins[0].res;
    if (n <= 2 && same) {
      int strips = (wins[0].res + kOutW - 1) / kOutW;
      chunk = fused_chunk_rows(wins[0].res, strips, n, resident);
    } else {
      long long work = 0;
      for (int k = 0; k < n; k++) work += (long long)((wins[k].res + kOutW - 1) / kOutW) * wins[k].res;
      long long target = 2LL * resident;
      chunk = (int)((work + target - 1) / target);
      if (chunk < 16) chunk = 16;
    }
  }
  out->exact = exact;
  out->full = full;
  out->project = project;
  out->wind = wind;
  out->chunk = chunk;
  out->nwin = n;
  out->descs.assign((size_t)n * sizeof(LayerArgs), 0);
  out->tile_window.clear();
  int tiles = 0;
  for (int k = 0; k < n; k++) {
    const FusedLayer& f = wins[k];
    LayerArgs L = make_layer(f.in, f.ho, f.uo, f.vo, f.a, f.boxes);
    L.hot.bed = f.in.bed;
    L.hot.res = f.res;
    L.hot.strips = (f.res + kOutW - 1) / kOutW;
    L.hot.chunk = chunk;
    L.hot.tile0 = tiles;
    L.hot.stats = f.stats;
    L.hot.totals = f.totals;
    int wt = L.hot.strips * ((f.res + chunk - 1) / chunk);
    if (n > 2) out->tile_window.insert(out->tile_window.end(), (size_t)wt, k);
    tiles += wt;
    memcpy(out->descs.data() + (size_t)k * sizeof(LayerArgs), &L, sizeof(LayerArgs));
  }
  out->tiles = tiles;
  out->rows.clear();
  if (xcd_w) {
    for (int x = 0; x < kXcds; x++)
      if (!(xcd_w[x] >= 0.25f && xcd_w[x] <= 4.0f)) return SILT_EINVAL;
    std::vector<LayerArgs> descs((size_t)n);
    memcpy(descs.data(), out->descs.data(), out->descs.size());
    out->rows = balanced_rows(descs, tiles, xcd_w);
  }
  return 0;
}

int fused_enqueue(const FusedLaunch& l, const void* dev_table, const int* dev_tile_window, hipStream_t st,
                  const int* dev_rows, unsigned long long* clocks, hipEvent_t ev_start, hipEvent_t ev_stop) {
  if (!dev_table || (l.nwin > 2 && !dev_tile_window) || l.tiles < 1) return SILT_EINVAL;
  FusedArgs fa{};
  fa.table = (ConstLayer*)dev_table;
  fa.tile_window = l.nwin > 2 ? (ConstInt*)dev_tile_window : nullptr;
  fa.rows = (ConstInt*)dev_rows;
  fa.clocks = clocks;
  fa.nwin = l.nwin;
  dispatch_any(&fa, l.tiles, l.exact, l.full, l.project, l.wind, st, ev_start, ev_stop);
  return hipGetLastError() == hipSuccess ? 0 : SILT_EDEVICE;
}

int fused_substep_windows(const FusedLayer* wins, int n, bool exact, int rows, hipStream_t st,
                          const FusedBalance* bal) {
  FusedLaunch l;
  int rc = fused_build(wins, n, exact, rows, &l, bal ? bal->xcd_w : nullptr);
  if (rc != 0) return rc;
  std::vector<LayerArgs> descs((size_t)n);
  memcpy(descs.data(), l.descs.data(), l.descs.size());
  FusedArgs fa{};
  rc = upload_descs(descs, l.tile_window, l.rows, st, &fa);
  if (rc != 0) return rc;
  return fused_enqueue(l, (const void*)fa.table, (const int*)fa.tile_window, st, (const int*)fa.rows,
                       bal ? bal->clocks : nullptr, bal ? bal->ev_start : nullptr, bal ? bal->ev_stop : nullptr);
}

int fused_substep_layers(const FusedLayer* layers, int nlayers, const float* bed, int res, bool exact,
                         hipStream_t st, const FusedBalance* bal) {
  if (nlayers < 1 || nlayers > 2) return SILT_EINVAL;
  FusedLayer w[2];
  for (int k = 0; k < nlayers; k++) {
    w[k] = layers[k];
    w[k].in.bed = bed;
    w[k].res = res;
  }
  return fused_substep_windows(w, nlayers, exact, 0, st, bal);
}

// One planned step of one layer: nsub fused substeps ping-ponging between
// the caller's planes and the workspace copies.  Plans the fused kernel does
// not cover (pace 3 slices, Gauss-Seidel sweeps, overlapping boxes) return
// SILT_EUNSUPPORTED so the caller can take the exact path.
int fused_step(const SheetPlanes& s, const StepPlan& plan, int res, float cell, float dt, const silt_drive& d,
               const silt_sheet_opts& o, bool gate_uniform, bool nu_uniform, float nu0, hipStream_t st,
               SheetTotals* book) {
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
    f.totals = book;
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
