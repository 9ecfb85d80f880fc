/*
 * oracle/mlp.c -- CPU restatement of the tire sub-grid network's forward and
 * backward passes (the network defined in include/silt_mlp.h).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle/sheet.c).  Compiled twice, FP32 and
 * FP64.  No counterpart in the reference: this restates the network's own
 * definition, so the GPU kernels' forward outputs and gradients can be
 * checked against it.  With bf16 = 1 it rounds at the same points as the GPU
 * (inputs, weights, hidden activations after GELU, and the output and hidden
 * gradients that feed the next products); with bf16 = 0 the FP64 build is
 * the exact function, which the tests pin by finite differences.
 *
 * params: w1[128][1536], b1[128], w2[128][128], b2[128], w3[8][128], b3[8].
 * Loss: sum over n x 8 of (y - t)^2 / (8 n).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL float
#define SUFFIX _f32
#endif

typedef REAL real;
#define PASTE_(a, b) a##b
#define PASTE(a, b) PASTE_(a, b)
#define F(name) PASTE(name, SUFFIX)

#if defined(REAL_IS_DOUBLE)
#define R_ERF erf
#define R_EXP exp
#else
#define R_ERF erff
#define R_EXP expf
#endif

enum { IN = 1536, H = 128, OUT = 8 };
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + H * IN)
#define OFF_W2 (OFF_B1 + H)
#define OFF_B2 (OFF_W2 + H * H)
#define OFF_W3 (OFF_B2 + H)
#define OFF_B3 (OFF_W3 + OUT * H)
#define NPARAM (OFF_B3 + OUT)

/* Round to the nearest BF16 value (ties to even), kept in real. */
static real bf(real v, int on) {
  if (!on) return v;
  float f = (float)v;
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return v;
  u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
  memcpy(&f, &u, 4);
  return (real)f;
}

static real gelu(real x) { return (real)0.5 * x * ((real)1 + R_ERF(x * (real)0.70710678118654752)); }
static real gelu_grad(real x) {
  real cdf = (real)0.5 * ((real)1 + R_ERF(x * (real)0.70710678118654752));
  real pdf = R_EXP((real)-0.5 * x * x) * (real)0.39894228040143268;
  return cdf + x * pdf;
}

/* One layer: z[n][no] = in[n][ni] . w[no][ni]^T + b; weights rounded when bf16. */
static void layer(const real* w, const real* b, const real* in, size_t n, int ni, int no, int bf16, real* z) {
  for (size_t s = 0; s < n; s++)
    for (int o = 0; o < no; o++) {
      real acc = 0;
      for (int i = 0; i < ni; i++) acc += bf(in[s * ni + i], bf16) * bf(w[(size_t)o * ni + i], bf16);
      z[s * no + o] = acc + b[o];
    }
}

/* Forward; z1, a1, z2, a2 (n x 128 each) may be NULL. */
void F(oracle_mlp_forward)(const real* p, const real* x, size_t n, int bf16, real* y, real* z1o, real* a1o,
                           real* z2o, real* a2o) {
  real* z1 = z1o ? z1o : (real*)malloc(n * H * sizeof(real));
  real* a1 = a1o ? a1o : (real*)malloc(n * H * sizeof(real));
  real* z2 = z2o ? z2o : (real*)malloc(n * H * sizeof(real));
  real* a2 = a2o ? a2o : (real*)malloc(n * H * sizeof(real));
  layer(p + OFF_W1, p + OFF_B1, x, n, IN, H, bf16, z1);
  for (size_t k = 0; k < n * H; k++) a1[k] = bf(gelu(z1[k]), bf16);
  layer(p + OFF_W2, p + OFF_B2, a1, n, H, H, bf16, z2);
  for (size_t k = 0; k < n * H; k++) a2[k] = bf(gelu(z2[k]), bf16);
  layer(p + OFF_W3, p + OFF_B3, a2, n, H, OUT, bf16, y);
  if (!z1o) free(z1);
  if (!a1o) free(a1);
  if (!z2o) free(z2);
  if (!a2o) free(a2);
}

/* Gradient of the loss into grad[NPARAM]; returns the sum of squared errors. */
double F(oracle_mlp_gradients)(const real* p, const real* x, const real* t, size_t n, int bf16, real* grad) {
  real* z1 = (real*)malloc(n * H * sizeof(real));
  real* a1 = (real*)malloc(n * H * sizeof(real));
  real* z2 = (real*)malloc(n * H * sizeof(real));
  real* a2 = (real*)malloc(n * H * sizeof(real));
  real* y = (real*)malloc(n * OUT * sizeof(real));
  real* dy = (real*)malloc(n * OUT * sizeof(real));
  real* dz2 = (real*)malloc(n * H * sizeof(real));
  real* dz1 = (real*)malloc(n * H * sizeof(real));
  F(oracle_mlp_forward)(p, x, n, bf16, y, z1, a1, z2, a2);
  memset(grad, 0, NPARAM * sizeof(real));
  double sse = 0.0;
  const real scale = (real)1 / (real)(n * OUT);
  for (size_t k = 0; k < n * OUT; k++) {
    real d = y[k] - t[k];
    sse += (double)d * (double)d;
    dy[k] = (real)2 * d * scale;
    grad[OFF_B3 + k % OUT] += dy[k];
  }
  /* dW3[o][j] = sum_s dy[s][o] a2[s][j];  dz2 = (dy W3) * gelu'(z2) */
  for (int o = 0; o < OUT; o++)
    for (int j = 0; j < H; j++) {
      real acc = 0;
      for (size_t s = 0; s < n; s++) acc += bf(dy[s * OUT + o], bf16) * a2[s * H + j];
      grad[OFF_W3 + o * H + j] = acc;
    }
  for (size_t s = 0; s < n; s++)
    for (int j = 0; j < H; j++) {
      real acc = 0;
      for (int o = 0; o < OUT; o++) acc += bf(dy[s * OUT + o], bf16) * bf(p[OFF_W3 + o * H + j], bf16);
      dz2[s * H + j] = acc * gelu_grad(z2[s * H + j]);
      grad[OFF_B2 + j] += dz2[s * H + j];
    }
  for (int o = 0; o < H; o++)
    for (int j = 0; j < H; j++) {
      real acc = 0;
      for (size_t s = 0; s < n; s++) acc += bf(dz2[s * H + o], bf16) * a1[s * H + j];
      grad[OFF_W2 + o * H + j] = acc;
    }
  for (size_t s = 0; s < n; s++)
    for (int j = 0; j < H; j++) {
      real acc = 0;
      for (int o = 0; o < H; o++) acc += bf(dz2[s * H + o], bf16) * bf(p[OFF_W2 + o * H + j], bf16);
      dz1[s * H + j] = acc * gelu_grad(z1[s * H + j]);
      grad[OFF_B1 + j] += dz1[s * H + j];
    }
  for (int o = 0; o < H; o++)
    for (int i = 0; i < IN; i++) {
      real acc = 0;
      for (size_t s = 0; s < n; s++) acc += bf(dz1[s * H + o], bf16) * bf(x[s * IN + i], bf16);
      grad[OFF_W1 + (size_t)o * IN + i] = acc;
    }
  free(z1);
  free(a1);
  free(z2);
  free(a2);
  free(y);
  free(dy);
  free(dz2);
  free(dz1);
  return sse;
}
