/*
 * silt_pair.h -- two-layer sheet handle of libsilt_amd.so.
 *
 * Water and slurry over one bed at any grid size, state resident on the GPU.
 * One silt_pair_step is one reference step of both layers: the generic-res
 * step_sheet (src/fluid_core.rs:770-785) on each layer with the layer
 * constants of FluidCore::step_pair (src/fluid_core.rs:8036-8125): water
 * Gravity{vorticity 0.08} at the water viscosity 0.9 Pa s
 * (dynamic_viscosity(Grain::Water, 1.0), :642-646), slurry
 * Gravity{vorticity 0.05} at its own viscosity plane, eps 1e-6, cap 8.
 * The ledger replaces ledger() (src/fluid_core.rs:6821-6883) for these two
 * layers, in FP64 with explicit dry-cut and transport-residue lines.
 * A handle is not thread-safe; calls on one handle are serialised on its
 * own HIP stream.
 */
#ifndef SILT_PAIR_H
#define SILT_PAIR_H

#include <stddef.h>
#include <stdint.h>

#include "silt_sheet.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct silt_pair silt_pair;

typedef struct silt_layer_ledger_t {
    double volume_m3;  /* sum of max(h, 0) * cell^2 now */
    double initial_m3; /* the same at the last upload */
    double cut_m3;     /* deleted by the dry cut since the upload */
    double residue_m3; /* added by FP32 rounding and the max(0) clamp in transport */
    double gap_m3;     /* volume - initial + cut - residue: zero up to FP64 summation */
} silt_layer_ledger_t;

typedef struct silt_pair_ledger_t {
    silt_layer_ledger_t water;
    silt_layer_ledger_t slurry;
    double total_m3; /* FluidLedger::total_m3 restricted to the two sheets */
} silt_pair_ledger_t;

/* A res x res pair at cell metres, kernel path SILT_PATH_*.  NULL on failure. */
silt_pair *silt_pair_create(size_t res, float cell, int path);
void silt_pair_destroy(silt_pair *p);

/* Host planes (res * res each) in; nu_slurry may be NULL with a finite
 * nu_slurry_uniform.  Resets the ledger. */
int silt_pair_upload(silt_pair *p, const float *bed, const float *water, const float *uw, const float *vw,
                     const float *slurry, const float *us, const float *vs, const float *nu_slurry,
                     float nu_slurry_uniform);
int silt_pair_download(silt_pair *p, float *water, float *uw, float *vw, float *slurry, float *us, float *vs);

/* One step of both layers.  Returns the larger substep count (0: both dry or
 * dt <= 0), or a negative SILT_E* code. */
int silt_pair_step(silt_pair *p, float dt);

int silt_pair_ledger(silt_pair *p, silt_pair_ledger_t *out);

/* GPU time of the last step's kernels, milliseconds (HIP events on the
 * handle's stream). */
float silt_pair_last_kernel_ms(const silt_pair *p);
/* Bit l set when layer l's last plan read its planes back to the host (a
 * 2 cm sheet, or more than 4096 separate wet bodies; other multi-body wet
 * sets are labelled on the GPU). */
int silt_pair_last_fallback(const silt_pair *p);

/* Autotune the fused kernel's rows per workgroup (silt_set_chunk_rows) for
 * this handle's side: every candidate (n of them in rows; NULL or n = 0 for
 * a built-in set around the default, 0 meaning the default itself) is timed
 * over `reps` steps of dt from the handle's current state, which is restored
 * after each candidate; the fastest is pinned.  best_rows / best_ms (mean
 * kernel time of a step) may be NULL.  Allocates one temporary copy of the
 * six planes.  No counterpart in the reference: a tuning step of the GPU
 * path, run once before a long simulation. */
int silt_pair_autotune(silt_pair *p, float dt, const int *rows, size_t n, int reps, int *best_rows,
                       float *best_ms);

#ifdef __cplusplus
}
#endif

#endif /* SILT_PAIR_H */
