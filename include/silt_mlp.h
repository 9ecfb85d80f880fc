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

#include "silt_core.h"

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
/* The same outputs from the one-launch inference kernel the tire steps use
 * (K summed in another order, so equal to silt_mlp_forward up to FP32
 * rounding).  No counterpart in the reference. */
int silt_mlp_infer(silt_mlp *m, const float *x, size_t n, float *y);
/* Gradient of the loss against targets t[n][8] into grad (SILT_MLP_PARAMS
 * values); *sse = sum of squared errors.  Weights unchanged. */
int silt_mlp_gradients(silt_mlp *m, const float *x, const float *t, size_t n, double *sse, float *grad);
/* One Adam step with learning rate lr on the batch; *sse before the step. */
int silt_mlp_train_batch(silt_mlp *m, const float *x, const float *t, size_t n, float lr, double *sse);

/* Training set and trainer of the network, on the GPU.
 *
 * A recorded step is one step of a FluidCore on the nested patch
 * (SILT_TIRE_NESTED) with its tires pushed; beside it a coarse-only twin
 * (SILT_TIRE_COARSE) steps from the same state.  The step's samples are the
 * cells of the network's region (the patch's cells grown by 2, at most
 * 20 x 20): inputs from the state before the step; targets the least-norm
 * edge fluxes (zero through the region's border) that turn the twin's depths
 * into the nested ones, in millimetres, per layer [E, W, N, S] (E = +x,
 * N = +z).  Input channels, each a 16 x 16 stencil around the cell (offsets
 * -8..+7, zero off the window): bed minus the cell's bed (x 10 per metre),
 * water and slurry depth (x 10 per metre), water u and v minus the tires'
 * mean velocity (x 0.125 s/m, zero on dry cells), and a nominal tire pressure
 * (3.5 kN on 0.30 m x the tire's width, in 100 kPa, by cell coverage). */
typedef struct silt_trainer silt_trainer;
silt_trainer *silt_trainer_create(size_t max_steps, unsigned seed);
void silt_trainer_destroy(silt_trainer *t);
/* Steps c like silt_core_step and records the step (when c is nested and has
 * tires and room is left).  Returns the recorded step count, 0 when nothing
 * was recorded, or a negative code. */
int silt_trainer_record(silt_trainer *t, silt_core *c, float dt, silt_spray_event *events, size_t cap,
                        size_t *nevents);
long silt_trainer_steps(const silt_trainer *t);
long silt_trainer_samples(const silt_trainer *t);
/* iterations of Adam on random batches drawn from the first train_steps
 * recorded steps (0: all); *last_mse = the last batch's mean squared error. */
int silt_trainer_fit(silt_trainer *t, size_t train_steps, int iterations, int batch, float lr, double *last_mse);
/* Mean squared error of the network over every sample of recorded steps
 * [first, first + count), and of predicting zero (the coarse-only step). */
int silt_trainer_eval(silt_trainer *t, size_t first, size_t count, double *mse_net, double *mse_zero);
int silt_trainer_params(const silt_trainer *t, float *params);
int silt_trainer_set_params(silt_trainer *t, const float *params);

/* Test hooks (host planes of the 200 x 200 window, region [i0, i0 + nx) x
 * [j0, j0 + nz), 2..20 cells a side).  targets_debug: the training targets
 * of a nested / coarse-only pair of depth planes ([cell][8], millimetres).
 * apply_debug: apply predicted fluxes y ([cell][8], millimetres) to water and
 * slurry in place, as a surrogate step does. */
int silt_surrogate_targets_debug(const float *hn_w, const float *hc_w, const float *hn_s, const float *hc_s, int i0,
                                 int j0, int nx, int nz, float *targets);
int silt_surrogate_apply_debug(float *water, float *slurry, const float *y, int i0, int j0, int nx, int nz);

#ifdef __cplusplus
}
#endif

#endif /* SILT_MLP_H */
