/*
 * silt_mlp.h -- the tire sub-grid network on its own (its forward, backward
 * and update kernels), for checking it against the CPU oracle and training it
 * outside a FluidCore.
 *
 * No counterpart in the reference, which always runs the nested 2 cm sheet
 * (src/fluid_core.rs:8316-8555): a learned stand-in for callers that want the
 * tire patch's effect on the coarse sheet at a fraction of its cost.
 *
 * A 3-layer MLP: 1536 inputs (a 16 x 16 stencil of coarse cells x 6
 * channels), two hidden layers of 128 with GELU (x/2 (1 + erf(x/sqrt 2))),
 * 8 outputs (4 edge fluxes x 2 layers).  Weights and activations are BF16
 * (round to nearest even), products accumulate in FP32 on the matrix cores;
 * FP32 master weights, Adam (0.9, 0.999, 1e-8).  Loss: mean squared error
 * over the n x 8 outputs.
 *
 * Parameter vector (FP32, SILT_MLP_PARAMS values, in order): w1[128][1536],
 * b1[128], w2[128][128], b2[128], w3[8][128], b3[8]; w[o][i] multiplies
 * input i of output o.
 */
#ifndef SILT_MLP_H
#define SILT_MLP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SILT_MLP_IN 1536
#define SILT_MLP_HIDDEN 128
#define SILT_MLP_OUT 8
#define SILT_MLP_PARAMS 214280

typedef struct silt_mlp silt_mlp;

/* Weights drawn from seed (uniform He initialisation, output layer x 0.1,
 * biases zero).  NULL on failure. */
silt_mlp *silt_mlp_create(unsigned seed);
void silt_mlp_destroy(silt_mlp *m);

int silt_mlp_get_params(const silt_mlp *m, float *params);
int silt_mlp_set_params(silt_mlp *m, const float *params);

/* y[n][8] = net(x[n][1536]) (host arrays). */
int silt_mlp_forward(silt_mlp *m, const float *x, size_t n, float *y);
/* Gradient of the loss against targets t[n][8] into grad (SILT_MLP_PARAMS
 * values); *sse = sum of squared errors.  Weights unchanged. */
int silt_mlp_gradients(silt_mlp *m, const float *x, const float *t, size_t n, double *sse, float *grad);
/* One Adam step with learning rate lr on the batch; *sse before the step. */
int silt_mlp_train_batch(silt_mlp *m, const float *x, const float *t, size_t n, float lr, double *sse);

#ifdef __cplusplus
}
#endif

#endif /* SILT_MLP_H */
