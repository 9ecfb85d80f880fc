// The tire sub-grid network: a 3-layer MLP (1536 -> 128 -> 128 -> 8, GELU)
// with BF16 weights and activations and FP32 accumulation on the MFMA units,
// trained with Adam on FP32 master weights.  Forward, backward and update
// are hand-written kernels (mlp.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace silt {

constexpr int kMlpIn = 1536;  // 16 x 16 coarse cells x 6 channels
constexpr int kMlpHidden = 128;
constexpr int kMlpOut = 8;  // 4 edge fluxes x 2 layers

// Parameter layout (FP32 master, gradients, Adam moments; the BF16 copies of
// the weights share the offsets).
constexpr size_t kOffW1 = 0;
constexpr size_t kOffB1 = kOffW1 + (size_t)kMlpHidden * kMlpIn;
constexpr size_t kOffW2 = kOffB1 + kMlpHidden;
constexpr size_t kOffB2 = kOffW2 + (size_t)kMlpHidden * kMlpHidden;
constexpr size_t kOffW3 = kOffB2 + kMlpHidden;
constexpr size_t kOffB3 = kOffW3 + (size_t)kMlpOut * kMlpHidden;
constexpr size_t kMlpParams = kOffB3 + kMlpOut;
constexpr int kMlpMaxSplits = 16;  // split-K slices of one product

struct Mlp {
  float* master = nullptr;  // [kMlpParams]
  uint16_t* wb = nullptr;   // [kMlpParams] BF16 weights (bias slots unused)
  float* grad = nullptr;
  float* m1 = nullptr;  // Adam first moment
  float* m2 = nullptr;  // Adam second moment
  int t = 0;            // Adam step
  // Activations for up to cap samples.
  size_t cap = 0;
  uint16_t* xb = nullptr;  // [cap x 1536]
  float* z1 = nullptr;     // [cap x 128] pre-activations
  uint16_t* a1 = nullptr;
  float* z2 = nullptr;
  uint16_t* a2 = nullptr;
  float* y = nullptr;  // [cap x 8]
  float* tg = nullptr;  // [cap x 8] targets
  float* dyf = nullptr;
  uint16_t* dy = nullptr;
  float* dz2f = nullptr;
  uint16_t* dz2 = nullptr;
  float* dz1f = nullptr;
  uint16_t* dz1 = nullptr;
  float* part = nullptr;   // split-K slices [kMlpMaxSplits][max(cap, 1536)][128]
  double* loss = nullptr;  // [1]
  hipStream_t st = nullptr;
};

// Allocates parameters and activations for cap samples; weights drawn from
// seed (uniform He initialisation, biases zero).  Returns 0 or a SILT_E* code.
int mlp_init(Mlp& m, size_t cap, unsigned seed, hipStream_t st);
int mlp_reserve(Mlp& m, size_t cap);
void mlp_free(Mlp& m);
// BF16 copies of the master weights (after a host write of the master).
void mlp_sync_bf16(Mlp& m);
// x (device, n x 1536 FP32) -> m.xb.
void mlp_load_x(Mlp& m, const float* x, size_t n);
// Forward of m.xb[0..n) into m.y (and the activations kept for backward).
void mlp_forward(Mlp& m, size_t n);
// Inference only: the network's outputs for m.xb[0..n) into m.y in one
// launch (K summed in another order than mlp_forward's; no activations
// kept for backward).
void mlp_infer(Mlp& m, size_t n);
// Backward of the mean squared error against m.tg over n samples after
// mlp_forward: gradients into m.grad, the sum of squared errors into m.loss.
void mlp_backward(Mlp& m, size_t n);
// One Adam step on the master weights with learning rate lr, then BF16 copies.
void mlp_adam(Mlp& m, float lr);

// FP32 <-> BF16, round to nearest even.
__host__ __device__ inline uint16_t f2bf(float f) {
  union {
    float f;
    uint32_t u;
  } c{f};
  if ((c.u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
  uint32_t r = c.u + 0x7fffu + ((c.u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b) {
  union {
    uint32_t u;
    float f;
  } c{(uint32_t)b << 16};
  return c.f;
}

}  // namespace silt
