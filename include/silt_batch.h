/*
 * silt_batch.h -- many independent sheet windows stepped together.
 *
 * Each window is one layer stepped as the reference's generic-res
 * step_sheet (src/fluid_core.rs:770-785) with a Gravity drive: its own side
 * res[k] and cell size, bed, depth and velocity planes, a uniform viscosity
 * (Pa s) and a vorticity gain.  One silt_batch_step advances every window by
 * one reference step; all windows' substeps of one index run in a single
 * launch of the fused kernel, the launches of a step replay from a captured
 * hipGraph while the plan's launch shape stays the same.
 *
 * No counterpart in the reference, which steps one window at a time (its
 * WINDOW_RES is fixed at 200, src/fluid_core.rs:72): a batching entry for
 * callers that run many small sheets (scenes, tiles of a large terrain).
 * Negative returns are SILT_E* codes (silt_sheet.h).
 */
#ifndef SILT_BATCH_H
#define SILT_BATCH_H

#include <stddef.h>

#include "silt_pair.h"
#include "silt_sheet.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct silt_batch silt_batch;

/* n windows, window k of res[k] x res[k] cells of cell[k] metres.  path:
 * SILT_PATH_FAST or SILT_PATH_FUSED_EXACT.  NULL on failure. */
silt_batch *silt_batch_create(size_t n, const size_t *res, const float *cell, int path);
void silt_batch_destroy(silt_batch *b);

/* Window k's state from the host (row-major res[k]^2 each), its viscosity
 * and vorticity gain.  Resets the window's ledger to this state. */
int silt_batch_upload(silt_batch *b, size_t k, const float *bed, const float *h, const float *u, const float *v,
                      float nu, float vorticity);
int silt_batch_download(silt_batch *b, size_t k, float *h, float *u, float *v);

/* One step of every window.  Returns the largest substep count (0: all dry
 * or dt <= 0) or a negative code. */
int silt_batch_step(silt_batch *b, float dt);

/* FP64 ledger of window k (the layer ledger of silt_pair.h). */
int silt_batch_ledger(silt_batch *b, size_t k, silt_layer_ledger_t *out);

/* GPU time of the last step (HIP events around its launches). */
float silt_batch_last_kernel_ms(const silt_batch *b);
/* Windows whose last plan needed a read back of their planes (more than
 * 4096 separate wet clusters); the launch still ran on the GPU. */
int silt_batch_last_fallbacks(const silt_batch *b);
/* Launch graphs captured so far (a new one each time the launch shape of a
 * step changes). */
int silt_batch_graphs_captured(const silt_batch *b);

/* Replay the step's launches from a captured hipGraph (default on). */
int silt_batch_use_graph(silt_batch *b, int on);
/* Rows per workgroup of the batch's launches (0: about two tiles per
 * resident workgroup slot over all windows). */
int silt_batch_set_rows(silt_batch *b, int rows);
/* Time every candidate rows-per-workgroup (n of them; NULL or n = 0: a
 * built-in set) over reps steps from the current state, which is restored
 * after each; keep the fastest.  With XCD weights given, the fastest height
 * is also timed with them switched the other way, and the faster setting
 * is kept.  best_rows / best_ms may be NULL. */
int silt_batch_autotune(silt_batch *b, float dt, const int *rows, size_t n, int reps, int *best_rows,
                        float *best_ms);

/* XCD weights for the batch's launches (as silt_pair_set_xcd_weights: the
 * rows of the tiles XCD x runs scaled by weights[x], in [0.25, 4]; NULL
 * clears them).  Given weights are on until silt_batch_autotune times the
 * best height with and without them and keeps the faster;
 * silt_batch_xcd_weights_on says which it kept.  Results do not depend on
 * them.  No counterpart in the reference: a tuning knob of the GPU path,
 * usually fed the weights silt_pair_balance calibrated on the same chip. */
int silt_batch_set_xcd_weights(silt_batch *b, const float *weights);
int silt_batch_xcd_weights_on(const silt_batch *b);

/* Test hook, no counterpart in the reference (whose sheet_work_boxes,
 * src/fluid_core.rs:1325-1447, it answers for device planes): the work boxes
 * of n host depth planes (h[k], res[k] x res[k], cell metres, wet above eps)
 * planned together on the GPU as a batch step plans them.  counts[k] = the
 * number of boxes of sheet k (boxes + 4 * cap * k holds them as i0, i1, j0,
 * j1), 0 for a dry sheet, -1 when the sheet would be planned on the host. */
int silt_plan_boxes_debug(const float *const *h, const int *res, size_t n, float cell, float eps, int *counts,
                          int *boxes, size_t cap);

#ifdef __cplusplus
}
#endif

#endif /* SILT_BATCH_H */
