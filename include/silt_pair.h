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

#include "silt_core.h"
#include "silt_rut.h"
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
 * dt <= 0), or a negative SILT_E* code.  The step's launches may still run
 * when it returns; download, ledger and silt_pair_last_kernel_ms wait for
 * them.  With speculation on (the default), the last step's plan is
 * launched before this step's plan is confirmed from the statistics; a
 * different plan undoes it (the starting planes are untouched, the ledger
 * lines restored) and launches the right one, so results do not depend on
 * it.  After an undone speculation the next steps plan before they launch,
 * for 1, 2, 4 ... at most 64 steps as misses repeat; one that holds resets
 * the pause. */
int silt_pair_step(silt_pair *p, float dt);

int silt_pair_ledger(silt_pair *p, silt_pair_ledger_t *out);

/* GPU time of the last step's kernels, milliseconds (HIP events on the
 * handle's stream; a step of one fused launch carries them in the launch's
 * own dispatch, so they time the kernel alone; waits for the step's
 * launches). */
float silt_pair_last_kernel_ms(const silt_pair *p);
/* Speculative launch of the last plan (default on); no counterpart in the
 * reference: a host-latency hiding switch of the GPU path.  Counts of
 * speculations that held and that were undone. */
int silt_pair_set_speculation(silt_pair *p, int on);
int silt_pair_speculation_counts(const silt_pair *p, long long *hits, long long *misses);
/* Bit l set when layer l's last plan read its planes back to the host (a
 * 2 cm sheet, or more than 4096 separate wet bodies; other multi-body wet
 * sets are labelled on the GPU). */
int silt_pair_last_fallback(const silt_pair *p);

/* Autotune the fused kernel's rows per workgroup (silt_set_chunk_rows) for
 * this handle's side: every candidate (n of them in rows; NULL or n = 0 for
 * a built-in set around the default, 0 meaning the default itself) is timed
 * over `reps` steps of dt from the handle's current state, which is restored
 * after each candidate; the fastest is pinned for launches that carry both
 * layers (a launch of one layer, a layer's extra substeps, keeps the
 * occupancy default).  best_rows / best_ms (mean
 * kernel time of a step) may be NULL.  Allocates one temporary copy of the
 * six planes.  No counterpart in the reference: a tuning step of the GPU
 * path, run once before a long simulation. */
int silt_pair_autotune(silt_pair *p, float dt, const int *rows, size_t n, int reps, int *best_rows,
                       float *best_ms);

/* XCD balance of the fused launch.  The 8 XCDs take the launch's workgroups
 * round robin, and on one chip the same work takes longer on some XCDs than
 * on others, so the launch waits for the slowest.  weights[x] scales the
 * rows of the tiles XCD x runs (each strip's rows are split in proportion to
 * the weights of the XCDs that run its chunks); NULL restores uniform
 * chunks.  Weights in [0.25, 4].  Results do not depend on the split.
 *
 * silt_pair_balance calibrates them on this chip.  Twelve rounds of reps
 * steps record every workgroup's start and end on the real-time clock:
 * round 0 with uniform chunks, each later one moving rows, from the best
 * round so far, toward the XCDs whose last workgroup ended early (a smaller
 * move after a round that was not faster) (the oldest workgroups on a CU are issued first, so
 * one XCD's tiles end at very different times: what counts is when the XCD
 * is done).  Each round is scored by the launch's mean kernel time and the
 * fastest weights are kept, uniform chunks included (ms[0] uniform, ms[1]
 * kept).  tile_us[0..7] is when each XCD finished with uniform chunks
 * (microseconds after the launch's first workgroup started), tile_us[8..15]
 * in the last round.
 * The steps are not the caller's: planes, ledger lines and speculation state
 * are restored.  weights, ms and tile_us may be NULL.  No counterpart in the
 * reference: a tuning step of the GPU path. */
int silt_pair_set_xcd_weights(silt_pair *p, const float *weights);
int silt_pair_balance(silt_pair *p, float dt, int reps, float *weights, float *ms, float *tile_us);
/* The per-workgroup records of the last calibration step (4 values each:
 * real-time clock at start and end at 100 MHz; pipeline rows | XCC_ID << 32;
 * tile | HW_ID << 32), at most n into out (may be NULL).  Returns how many
 * there are (workgroups of the largest launch the handle can make; unused
 * ones are zero). */
size_t silt_pair_clock_records(const silt_pair *p, unsigned long long *out, size_t n);

/* Tires on the sheet.  No counterpart in the reference, whose tires act on
 * the 40 m FluidCore window only (FluidCore::push_tire and the rut stamp,
 * src/fluid_core.rs:5964, src/soil.rs:2526-2879): the same coupling on a
 * sheet of any size, for callers that drive many vehicles over a large one.
 *
 * set_ground: the ground under the sheet (no placed soil) and the world
 * corner of cell (0, 0) (default: the sheet centred on the origin); the sheet
 * bed becomes ground bed + placed.  get_ground: bed, placed and rut depth
 * (initial - bed, >= 0) out; any may be NULL. */
int silt_pair_set_ground(silt_pair *p, const float *bed, float origin_x, float origin_z);
int silt_pair_get_ground(silt_pair *p, float *bed, float *placed, float *rut);
/* The tire sub-grid network (SILT_MLP_PARAMS floats, silt_mlp.h) each
 * vehicle of silt_pair_step_tires uses; NULL: none (the tires only shove). */
int silt_pair_set_surrogate(silt_pair *p, const float *params);
/* One step with vehicles: the contacts cut Bekker ruts into the ground (when
 * set) and the bed follows; the network reads each vehicle's region before
 * the step (pushes in groups of per_vehicle, one vehicle each); the pushes
 * shove both layers (push_on, src/fluid_core.rs:8152-8239); both layers step
 * as silt_pair_step; the network's edge fluxes are applied.  When there are
 * as many contacts as pushes, they come in the same vehicle groups, and
 * vehicles' rut footprints must not overlap (each vehicle's ruts are cut by
 * its own workgroup; a vehicle's own contacts stay in order).  Returns as
 * silt_pair_step. */
int silt_pair_step_tires(silt_pair *p, float dt, const silt_tire_push *pushes, size_t npush, size_t per_vehicle,
                         const silt_tire_contact *contacts, size_t ncontacts);
/* GPU time of the last silt_pair_step_tires' tire work (network inputs and
 * forward, shoves, flux apply), milliseconds; 0 before a network is set. */
float silt_pair_last_tire_ms(const silt_pair *p);
/* Launch shape of the tire work, no counterpart in the reference: on (the
 * default), vehicles whose network regions and push boxes do not overlap
 * run each stage in one launch; off, one vehicle after another.  The results
 * are the same either way. */
int silt_pair_set_tire_batching(silt_pair *p, int on);

#ifdef __cplusplus
}
#endif

#endif /* SILT_PAIR_H */
