// MFMA kernels of the tire sub-grid network (mlp.hpp).
//
// One GEMM kernel serves every product of the forward and backward passes:
// C(m, n) = sum_k A(m, k) B(n, k) with BF16 operands, each stored either
// k-contiguous or transposed, and FP32 accumulation in
// v_mfma_f32_16x16x16_bf16.  A 256-lane workgroup owns a 64 x 64 tile of C;
// each of its four wavefronts owns 32 x 32 (2 x 2 MFMA tiles) and the K
// dimension streams through LDS 32 at a time.  The epilogue is fused:
//   EPI_PLAIN     C (+ bias)                          -> FP32
//   EPI_HIDDEN    z = C + bias -> FP32, gelu(z)       -> BF16
//   EPI_GELU_BWD  dz = C * gelu'(z)                   -> FP32 and BF16
//   EPI_PARTIAL   one K slice of C (split K)          -> FP32 slice, summed
//                 in order by the next kernel
// A product with few output tiles (the first layer at inference sizes, the
// first layer's weight gradient) splits K over gridDim.z to fill the chip.
//
// Forward (n samples):  z1 = x W1^T + b1, a1 = gelu(z1); z2 = a1 W2^T + b2,
// a2 = gelu(z2); y = a2 W3^T + b3.  Backward of L = sum (y - t)^2 / (8 n):
// dy = 2 (y - t) / (8 n); dW3 = dy^T a2; dz2 = (dy W3) * gelu'(z2);
// dW2 = dz2^T a1; dz1 = (dz2 W2) * gelu'(z1); dW1 = dz1^T x; db = column sums.
#include <math.h>
#include <string.h>

#include <vector>

#include "mlp.hpp"
#include "silt_sheet.h"

namespace silt {

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

enum { EPI_PLAIN = 0, EPI_HIDDEN = 1, EPI_GELU_BWD = 2, EPI_PARTIAL = 3 };

struct Gemm {
  int M, N, K;
  const uint16_t* A;
  int lda;
  const uint16_t* B;
  int ldb;
  const float* bias;
  float* out;
  int ldo;
  uint16_t* outb;
  int ldob;
  const float* z;
  int ldz;
  int kchunk;  // K per workgroup along gridDim.z (split K, EPI_PARTIAL); 0: all of K
};

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x) {
  float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  float pdf = expf(-0.5f * x * x) * 0.39894228040143268f;
  return cdf + x * pdf;
}

constexpr int kT = 64;    // tile side
constexpr int kKs = 32;   // K per LDS stage
constexpr int kLd = 40;   // LDS row stride (elements): 80 B keeps 8 B alignment

// AT: A stored as [K][lda] (m contiguous); otherwise [M][lda] (k contiguous).
// BT: B stored as [K][ldb] (n contiguous); otherwise [N][ldb].
template <bool AT, bool BT, int EPI>
__global__ __launch_bounds__(256) void k_gemm(Gemm g) {
  __shared__ __attribute__((aligned(16))) uint16_t As[kT][kLd];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[kT][kLd];
  const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = (int)blockIdx.y * kT, n0 = (int)blockIdx.x * kT;
  v4f acc[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
  const int kb = g.kchunk ? (int)blockIdx.z * g.kchunk : 0;
  const int ke = g.kchunk ? min(g.K, kb + g.kchunk) : g.K;
  for (int k0 = kb; k0 < ke; k0 += kKs) {
    // Stage the A and B tiles, reading along each operand's contiguous axis:
    // a k-contiguous operand moves 8 values (16 B) per lane when the row is
    // aligned and the chunk lies inside K.
    if (!AT) {
      const int r = tid >> 2, c = (tid & 3) * 8;
      const int m = m0 + r, k = k0 + c;
      if (m < g.M && k + 8 <= ke && (g.lda & 7) == 0) {
        *reinterpret_cast<uint4*>(&As[r][c]) = *reinterpret_cast<const uint4*>(g.A + (size_t)m * g.lda + k);
      } else {
        for (int q = 0; q < 8; q++) As[r][c + q] = (m < g.M && k + q < ke) ? g.A[(size_t)m * g.lda + k + q] : 0;
      }
    } else {
      for (int e = tid; e < kT * kKs; e += 256) {
        const int c = e / kT, r = e % kT;
        const int m = m0 + r, k = k0 + c;
        As[r][c] = (m < g.M && k < ke) ? g.A[(size_t)k * g.lda + m] : (uint16_t)0;
      }
    }
    if (!BT) {
      const int r = tid >> 2, c = (tid & 3) * 8;
      const int n = n0 + r, k = k0 + c;
      if (n < g.N && k + 8 <= ke && (g.ldb & 7) == 0) {
        *reinterpret_cast<uint4*>(&Bs[r][c]) = *reinterpret_cast<const uint4*>(g.B + (size_t)n * g.ldb + k);
      } else {
        for (int q = 0; q < 8; q++) Bs[r][c + q] = (n < g.N && k + q < ke) ? g.B[(size_t)n * g.ldb + k + q] : 0;
      }
    } else {
      for (int e = tid; e < kT * kKs; e += 256) {
        const int c = e / kT, r = e % kT;
        const int n = n0 + r, k = k0 + c;
        Bs[r][c] = (n < g.N && k < ke) ? g.B[(size_t)k * g.ldb + n] : (uint16_t)0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kKs; ks += 16) {
      // Lane l holds rows (l mod 16) of its 16-row slices, k = 4 (l / 16) + 0..3.
      const int kk = ks + 4 * (lane >> 4);
      v4s a[2], b[2];
#pragma unroll
      for (int t = 0; t < 2; t++) {
        a[t] = *reinterpret_cast<const v4s*>(&As[wm * 32 + t * 16 + (lane & 15)][kk]);
        b[t] = *reinterpret_cast<const v4s*>(&Bs[wn * 32 + t * 16 + (lane & 15)][kk]);
      }
#pragma unroll
      for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
          acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
    }
    __syncthreads();
  }
  // Result layout: lane l holds C[4 (l / 16) + i][l mod 16] of each 16 x 16 tile.
#pragma unroll
  for (int ti = 0; ti < 2; ti++)
#pragma unroll
    for (int tj = 0; tj < 2; tj++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int m = m0 + wm * 32 + ti * 16 + 4 * (lane >> 4) + i;
        const int n = n0 + wn * 32 + tj * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        float v = acc[ti][tj][i];
        if (EPI == EPI_PARTIAL) {
          // Slice blockIdx.z of a [splits][M][ldo] buffer, summed in a fixed
          // order afterwards (deterministic, unlike atomics).
          g.out[((size_t)blockIdx.z * g.M + m) * g.ldo + n] = v;
        } else if (EPI == EPI_PLAIN) {
          if (g.bias) v += g.bias[n];
          g.out[(size_t)m * g.ldo + n] = v;
        } else if (EPI == EPI_HIDDEN) {
          v += g.bias[n];
          g.out[(size_t)m * g.ldo + n] = v;
          g.outb[(size_t)m * g.ldob + n] = f2bf(gelu(v));
        } else {
          const float d = v * gelu_grad(g.z[(size_t)m * g.ldz + n]);
          g.out[(size_t)m * g.ldo + n] = d;
          g.outb[(size_t)m * g.ldob + n] = f2bf(d);
        }
      }
}

template <bool AT, bool BT, int EPI>
void gemm(const Gemm& g, hipStream_t st) {
  const unsigned splits = g.kchunk ? (unsigned)((g.K + g.kchunk - 1) / g.kchunk) : 1u;
  dim3 grid((unsigned)((g.N + kT - 1) / kT), (unsigned)((g.M + kT - 1) / kT), splits);
  hipLaunchKernelGGL((k_gemm<AT, BT, EPI>), grid, dim3(256), 0, st, g);
}

// Sum of the split-K slices in order, plus bias, into z (FP32) and gelu(z)
// (BF16): the epilogue of a split hidden layer.
__global__ void k_split_bias_gelu(const float* __restrict__ part, int splits, int M, int N,
                                  const float* __restrict__ bias, float* __restrict__ z, uint16_t* __restrict__ a) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t cnt = (size_t)M * N;
  if (i >= cnt) return;
  float v = 0.0f;
  for (int s = 0; s < splits; s++) v += part[(size_t)s * cnt + i];
  v += bias[i % N];
  z[i] = v;
  a[i] = f2bf(gelu(v));
}

// Sum of the split-K slices in order into out.
__global__ void k_split_sum(const float* __restrict__ part, int splits, size_t cnt, float* __restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  float v = 0.0f;
  for (int s = 0; s < splits; s++) v += part[(size_t)s * cnt + i];
  out[i] = v;
}

// Split K so a product with few output tiles still fills the chip: about 256
// workgroups, at most kMaxSplits slices of at least 64.
int split_chunk(int M, int N, int K) {
  const int tiles = ((M + kT - 1) / kT) * ((N + kT - 1) / kT);
  int splits = 256 / (tiles > 0 ? tiles : 1);
  if (splits > kMlpMaxSplits) splits = kMlpMaxSplits;
  if (splits < 2 || K < 128) return 0;
  int chunk = (K + splits - 1) / splits;
  chunk = ((chunk + kKs - 1) / kKs) * kKs;
  return chunk < 64 ? 64 : chunk;
}

__global__ void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// dy = 2 (y - t) / (8 n) (FP32 and BF16); sum of squared errors in FP64.
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ y, const float* __restrict__ t, size_t count,
                                              float scale, float* __restrict__ dyf, uint16_t* __restrict__ dy,
                                              double* loss) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (i < count) {
    const float d = y[i] - t[i];
    const float g = 2.0f * d * scale;
    dyf[i] = g;
    dy[i] = f2bf(g);
    s = (double)d * (double)d;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0 && s != 0.0) atomicAdd(loss, s);
}

// Bias gradient: column sums of an M x N FP32 matrix.  A block takes 32
// columns; its 8 row groups each sum every 8th row (coalesced 128 B rows),
// then the groups add in a fixed order.
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ x, int M, int N, int ld,
                                                float* __restrict__ out) {
  __shared__ float part[8][32];
  const int c = (int)(threadIdx.x & 31), g = (int)(threadIdx.x >> 5);
  const int n = (int)blockIdx.x * 32 + c;
  float s = 0.0f;
  if (n < N)
    for (int m = g; m < M; m += 8) s += x[(size_t)m * ld + n];
  part[g][c] = s;
  __syncthreads();
  if (g == 0 && n < N) {
    float t = part[0][c];
    for (int k = 1; k < 8; k++) t += part[k][c];
    out[n] = t;
  }
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, uint16_t* __restrict__ wb, size_t n, float lr, float c1, float c2) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float b1 = 0.9f, b2 = 0.999f;
  float gi = g[i];
  float a = b1 * m1[i] + (1.0f - b1) * gi;
  float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  float wi = w[i] - lr * (a * c1) / (sqrtf(b * c2) + 1.0e-8f);
  w[i] = wi;
  wb[i] = f2bf(wi);
}

__global__ void k_zero_f64(double* p) { *p = 0.0; }

template <typename T>
bool dalloc(T** p, size_t n) {
  return hipMalloc(p, n * sizeof(T)) == hipSuccess;
}

}  // namespace

int mlp_reserve(Mlp& m, size_t cap) {
  if (cap <= m.cap) return 0;
  void* old[] = {m.xb, m.z1, m.a1, m.z2, m.a2, m.y, m.tg, m.dyf, m.dy, m.dz2f, m.dz2, m.dz1f, m.dz1, m.part};
  if (m.st) (void)hipStreamSynchronize(m.st);
  for (void* p : old)
    if (p) (void)hipFree(p);
  m.cap = 0;
  const size_t H = kMlpHidden, O = kMlpOut;
  bool ok = dalloc(&m.xb, cap * kMlpIn) && dalloc(&m.z1, cap * H) && dalloc(&m.a1, cap * H) &&
            dalloc(&m.z2, cap * H) && dalloc(&m.a2, cap * H) && dalloc(&m.y, cap * O) && dalloc(&m.tg, cap * O) &&
            dalloc(&m.dyf, cap * O) && dalloc(&m.dy, cap * O) && dalloc(&m.dz2f, cap * H) &&
            dalloc(&m.dz2, cap * H) && dalloc(&m.dz1f, cap * H) && dalloc(&m.dz1, cap * H) &&
            dalloc(&m.part, (size_t)kMlpMaxSplits * (cap > (size_t)kMlpIn ? cap : (size_t)kMlpIn) * H);
  if (!ok) return SILT_ENOMEM;
  m.cap = cap;
  return 0;
}

int mlp_init(Mlp& m, size_t cap, unsigned seed, hipStream_t st) {
  m.st = st;
  if (!dalloc(&m.master, kMlpParams) || !dalloc(&m.wb, kMlpParams) || !dalloc(&m.grad, kMlpParams) ||
      !dalloc(&m.m1, kMlpParams) || !dalloc(&m.m2, kMlpParams) || !dalloc(&m.loss, 1))
    return SILT_ENOMEM;
  if (hipMemset(m.m1, 0, kMlpParams * 4) != hipSuccess || hipMemset(m.m2, 0, kMlpParams * 4) != hipSuccess ||
      hipMemset(m.grad, 0, kMlpParams * 4) != hipSuccess)
    return SILT_EDEVICE;
  // Uniform He initialisation from a splitmix64 stream; the output layer is
  // scaled down so an untrained network predicts small fluxes.
  std::vector<float> w(kMlpParams, 0.0f);
  uint64_t s = 0x9e3779b97f4a7c15ull ^ (uint64_t)seed;
  auto next = [&]() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (float)((double)(z >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f;
  };
  auto fill = [&](size_t off, size_t count, int fan_in, float gain) {
    float a = gain * sqrtf(6.0f / (float)fan_in);
    for (size_t i = 0; i < count; i++) w[off + i] = a * next();
  };
  fill(kOffW1, (size_t)kMlpHidden * kMlpIn, kMlpIn, 1.0f);
  fill(kOffW2, (size_t)kMlpHidden * kMlpHidden, kMlpHidden, 1.0f);
  fill(kOffW3, (size_t)kMlpOut * kMlpHidden, kMlpHidden, 0.1f);
  if (hipMemcpy(m.master, w.data(), kMlpParams * 4, hipMemcpyHostToDevice) != hipSuccess) return SILT_EDEVICE;
  mlp_sync_bf16(m);
  m.t = 0;
  return mlp_reserve(m, cap);
}

void mlp_free(Mlp& m) {
  if (m.st) (void)hipStreamSynchronize(m.st);
  void* ps[] = {m.master, m.wb, m.grad, m.m1, m.m2, m.loss, m.xb, m.z1, m.a1, m.z2, m.a2,
                m.y, m.tg, m.dyf, m.dy, m.dz2f, m.dz2, m.dz1f, m.dz1, m.part};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  hipStream_t st = m.st;
  m = Mlp{};
  m.st = st;
}

void mlp_sync_bf16(Mlp& m) {
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((kMlpParams + 255) / 256)), dim3(256), 0, m.st, m.master, m.wb,
                     kMlpParams);
}

void mlp_load_x(Mlp& m, const float* x, size_t n) {
  size_t c = n * kMlpIn;
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((c + 255) / 256)), dim3(256), 0, m.st, x, m.xb, c);
}

// Inference of the whole network for up to 16 rows per workgroup of 16
// wavefronts, in one launch.  Layer 1 (K = 1536) is split over the waves:
// wave w owns hidden columns 32 (w mod 4) .. +31 over the K quarter w / 4,
// reading the inputs and the weights straight from global memory (each lane
// reads 8 consecutive k of its row, split over two MFMA steps, so the
// product sums K in a permuted order); the quarters' partial sums meet in
// LDS and add in order.  The hidden activations stay in LDS; layer 2 runs
// on the first four waves, layer 3 on the VALU.
__global__ __launch_bounds__(1024) void k_mlp_infer(const uint16_t* __restrict__ x, const uint16_t* __restrict__ wb,
                                                    const float* __restrict__ master, float* __restrict__ y, int n) {
  __shared__ __attribute__((aligned(16))) uint16_t hs[16][kMlpHidden + 8];
  __shared__ float part[4][16][kMlpHidden];
  __shared__ uint16_t w3s[kMlpOut * kMlpHidden];
  const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // The output layer's weights (2 KB) into LDS while layer 1 runs.
  for (int e = tid; e < kMlpOut * kMlpHidden; e += 1024) w3s[e] = wb[kOffW3 + e];
  const int cg = wave & 3, kq = wave >> 2;
  const int r0 = (int)blockIdx.x * 16;
  const int row = r0 + (lane & 15);
  const bool row_in = row < n;
  const int q = lane >> 4;
  // Product of 16 rows of a bf16 operand A with weights W ([H][K], K
  // contiguous) over k in [kb, ke), for hidden columns 32 cg .. +31: lane
  // holds C[4 q + i][lane & 15] of each 16 x 16 tile.  K in groups of 128,
  // the loads of a group issued before its products.
  auto hidden = [&](auto loadA, const uint16_t* W, int K, int kb, int ke, v4f (&acc)[2]) {
    acc[0] = acc[1] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t* w0 = W + (size_t)(cg * 32 + (lane & 15)) * K;
    const uint16_t* w1 = w0 + (size_t)16 * K;
    for (int k0 = kb; k0 < ke; k0 += 128) {
      uint4 av[4], bv[4][2];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int k = k0 + 32 * u + 8 * q;
        av[u] = loadA(k);
        bv[u][0] = *reinterpret_cast<const uint4*>(w0 + k);
        bv[u][1] = *reinterpret_cast<const uint4*>(w1 + k);
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        v4s a0, a1;
        memcpy(&a0, &av[u], 8);
        memcpy(&a1, reinterpret_cast<const char*>(&av[u]) + 8, 8);
#pragma unroll
        for (int t = 0; t < 2; t++) {
          v4s b0, b1;
          memcpy(&b0, &bv[u][t], 8);
          memcpy(&b1, reinterpret_cast<const char*>(&bv[u][t]) + 8, 8);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b0, acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b1, acc[t], 0, 0, 0);
        }
      }
    }
  };
  v4f acc[2];
  constexpr int kQuarter = kMlpIn / 4;  // 384, a multiple of 128
  hidden(
      [&](int k) -> uint4 {
        return row_in ? *reinterpret_cast<const uint4*>(x + (size_t)row * kMlpIn + k) : uint4{0u, 0u, 0u, 0u};
      },
      wb + kOffW1, kMlpIn, kq * kQuarter, (kq + 1) * kQuarter, acc);
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int i = 0; i < 4; i++) part[kq][4 * q + i][cg * 32 + t * 16 + (lane & 15)] = acc[t][i];
  __syncthreads();
  // Quarters in order, bias, GELU: one hidden value per lane pair.
  for (int e = tid; e < 16 * kMlpHidden; e += 1024) {
    const int r = e / kMlpHidden, c = e % kMlpHidden;
    const float z = ((part[0][r][c] + part[1][r][c]) + part[2][r][c]) + part[3][r][c];
    hs[r][c] = f2bf(gelu(z + master[kOffB1 + c]));
  }
  __syncthreads();
  if (kq == 0) {
    hidden([&](int k) -> uint4 { return *reinterpret_cast<const uint4*>(&hs[lane & 15][k]); }, wb + kOffW2, kMlpHidden,
           0, kMlpHidden, acc);
  }
  __syncthreads();  // every wave has read layer 1's activations
  if (kq == 0) {
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int col = cg * 32 + t * 16 + (lane & 15);
        hs[4 * q + i][col] = f2bf(gelu(acc[t][i] + master[kOffB2 + col]));
      }
  }
  __syncthreads();
  // Output layer: lane t < 128 computes output (t / 8, t % 8).
  if (tid < 16 * kMlpOut) {
    const int r = tid / kMlpOut, o = tid % kMlpOut;
    float s = 0.0f;
    const uint16_t* w3 = w3s + o * kMlpHidden;
#pragma unroll 16
    for (int k = 0; k < kMlpHidden; k++) s += bf2f(hs[r][k]) * bf2f(w3[k]);
    if (r0 + r < n) y[(size_t)(r0 + r) * kMlpOut + o] = s + master[kOffB3 + o];
  }
}

void mlp_infer(Mlp& m, size_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + 15) / 16)), dim3(1024), 0, m.st, m.xb, m.wb, m.master, m.y,
                     (int)n);
}

void mlp_forward(Mlp& m, size_t n) {
  const int N = (int)n, H = kMlpHidden;
  Gemm g{};
  const int c1 = split_chunk(N, H, kMlpIn);
  if (c1) {
    const int splits = (kMlpIn + c1 - 1) / c1;
    g = Gemm{N, H, kMlpIn, m.xb, kMlpIn, m.wb + kOffW1, kMlpIn, nullptr, m.part, H, nullptr, 0, nullptr, 0, c1};
    gemm<false, false, EPI_PARTIAL>(g, m.st);
    const size_t cnt = (size_t)N * H;
    hipLaunchKernelGGL(k_split_bias_gelu, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, m.st, m.part, splits, N,
                       H, m.master + kOffB1, m.z1, m.a1);
  } else {
    g = Gemm{N, H, kMlpIn, m.xb, kMlpIn, m.wb + kOffW1, kMlpIn, m.master + kOffB1, m.z1, H, m.a1, H, nullptr, 0, 0};
    gemm<false, false, EPI_HIDDEN>(g, m.st);
  }
  g = Gemm{N, H, H, m.a1, H, m.wb + kOffW2, H, m.master + kOffB2, m.z2, H, m.a2, H, nullptr, 0, 0};
  gemm<false, false, EPI_HIDDEN>(g, m.st);
  g = Gemm{N, kMlpOut, H, m.a2, H, m.wb + kOffW3, H, m.master + kOffB3, m.y, kMlpOut, nullptr, 0, nullptr, 0, 0};
  gemm<false, false, EPI_PLAIN>(g, m.st);
}

namespace {

// A weight gradient dW = dZ^T A (both operands batch-major, K = batch): few
// output tiles and a long K, so K is split into slices summed in order.
void weight_grad(Gemm g, Mlp& m) {
  const int c = split_chunk(g.M, g.N, g.K);
  if (!c) {
    gemm<true, true, EPI_PLAIN>(g, m.st);
    return;
  }
  float* out = g.out;
  const int ldo = g.ldo;
  const int splits = (g.K + c - 1) / c;
  g.out = m.part;
  g.ldo = g.N;
  g.kchunk = c;
  gemm<true, true, EPI_PARTIAL>(g, m.st);
  const size_t cnt = (size_t)g.M * g.N;
  // The gradient rows are contiguous (ldo == N) for every weight matrix.
  (void)ldo;
  hipLaunchKernelGGL(k_split_sum, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, m.st, m.part, splits, cnt, out);
}

}  // namespace

void mlp_backward(Mlp& m, size_t n) {
  const int N = (int)n, H = kMlpHidden, O = kMlpOut;
  const size_t cnt = n * O;
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(1), 0, m.st, m.loss);
  hipLaunchKernelGGL(k_loss, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, m.st, m.y, m.tg, cnt,
                     1.0f / (float)cnt, m.dyf, m.dy, m.loss);
  Gemm g{};
  // dW3 = dy^T a2  (8 x 128, K = n)
  weight_grad(Gemm{O, H, N, m.dy, O, m.a2, H, nullptr, m.grad + kOffW3, H, nullptr, 0, nullptr, 0, 0}, m);
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(256), 0, m.st, m.dyf, N, O, O, m.grad + kOffB3);
  // dz2 = (dy W3) * gelu'(z2)  (n x 128, K = 8)
  g = Gemm{N, H, O, m.dy, O, m.wb + kOffW3, H, nullptr, m.dz2f, H, m.dz2, H, m.z2, H, 0};
  gemm<false, true, EPI_GELU_BWD>(g, m.st);
  // dW2 = dz2^T a1
  weight_grad(Gemm{H, H, N, m.dz2, H, m.a1, H, nullptr, m.grad + kOffW2, H, nullptr, 0, nullptr, 0, 0}, m);
  hipLaunchKernelGGL(k_colsum, dim3(4), dim3(256), 0, m.st, m.dz2f, N, H, H, m.grad + kOffB2);
  // dz1 = (dz2 W2) * gelu'(z1)
  g = Gemm{N, H, H, m.dz2, H, m.wb + kOffW2, H, nullptr, m.dz1f, H, m.dz1, H, m.z1, H, 0};
  gemm<false, true, EPI_GELU_BWD>(g, m.st);
  // dW1 = dz1^T x  (128 x 1536)
  weight_grad(Gemm{H, kMlpIn, N, m.dz1, H, m.xb, kMlpIn, nullptr, m.grad + kOffW1, kMlpIn, nullptr, 0, nullptr, 0, 0},
              m);
  hipLaunchKernelGGL(k_colsum, dim3(4), dim3(256), 0, m.st, m.dz1f, N, H, H, m.grad + kOffB1);
}

void mlp_adam(Mlp& m, float lr) {
  m.t++;
  const float c1 = 1.0f / (1.0f - powf(0.9f, (float)m.t));
  const float c2 = 1.0f / (1.0f - powf(0.999f, (float)m.t));
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((kMlpParams + 255) / 256)), dim3(256), 0, m.st, m.master, m.grad, m.m1,
                     m.m2, m.wb, kMlpParams, lr, c1, c2);
}

}  // namespace silt
