/*
 * silt_sheet.h -- C ABI of the MI355X sheet step (libsilt_amd.so).
 *
 * Drop-in boundary for the two-layer shallow-water sheet of x3-team/Dirty-gen
 * ("silt").  Plain pointers and sizes only.  Each entry point names the
 * reference interface it replaces (file:line inside the reference tree).
 *
 * Conventions
 *  - Planes are row-major, idx = j * res + i, FP32 (u8 for gate), as the
 *    reference's slices.
 *  - Invalid *values* are a silent no-op, exactly as in the reference
 *    (dt <= 0, res == 0, slices shorter than res * res, a dry sheet):
 *    the call returns 0 and leaves the planes untouched.
 *  - Invalid *arguments* the reference cannot express (NULL pointers, an
 *    unknown drive kind or option) return a negative SILT_E* code.
 *  - A non-negative return is the number of substeps taken.
 *  - Calls are not thread-safe against each other on the same buffers or
 *    handle; the *_device variants are asynchronous on the given HIP stream
 *    (NULL = the null stream).
 */
#ifndef SILT_SHEET_H
#define SILT_SHEET_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SILT_OK 0
#define SILT_EINVAL (-1)
#define SILT_EDEVICE (-2)
#define SILT_ENOMEM (-3)
#define SILT_EUNSUPPORTED (-4)
/* A size beyond the fixed maximum of an entry point that allocates nothing. */
#define SILT_ETOOBIG (-5)

/* SheetDrive, src/fluid_core.rs:204-211. */
#define SILT_DRIVE_GRAVITY 0
#define SILT_DRIVE_WIND 1
typedef struct silt_drive {
    int32_t kind;    /* SILT_DRIVE_GRAVITY or SILT_DRIVE_WIND */
    float x, z;      /* wind velocity (wind drive only), m/s */
    float vorticity; /* confinement strength */
} silt_drive;

/* Confinement schedule. */
#define SILT_CONFINE_REFERENCE 0 /* the reference's rule: Gauss-Seidel sweep on small
                                    coarse boxes and beside the other layer,
                                    snapshot Jacobi on the 2 cm sheet and big boxes */
#define SILT_CONFINE_JACOBI 1    /* snapshot Jacobi everywhere (fully parallel) */

/* Kernel path. */
#define SILT_PATH_EXACT 0 /* stage kernels, reference operation order, no FMA */
#define SILT_PATH_FAST 1  /* fused streaming kernel (FMA, reciprocals) */
#define SILT_PATH_FUSED_EXACT 2 /* fused streaming kernel, reference operation order */

/* The arguments of step_sheet_eps beyond step_sheet's
 * (src/fluid_core.rs:856-889, 821-854). */
typedef struct silt_sheet_opts {
    float eps;         /* dry threshold; step_sheet uses 1e-6 */
    int32_t banks;     /* dry-bank test in the open transport (drain_over_bed) */
    float sub_cap;     /* substep cap, clamped to [1, 8]; step_sheet uses 8 */
    int32_t pace;      /* 0..3, the 2 cm sheet's substep policy */
    int32_t row_split; /* 1: a plain step_sheet call (row splits allowed);
                          0: one layer of FluidCore::step_pair (serial_inner) */
    int32_t confine;   /* SILT_CONFINE_* */
    int32_t path;      /* SILT_PATH_* */
} silt_sheet_opts;

/* Fills the step_sheet defaults: eps 1e-6, no banks, cap 8, pace 0,
 * row_split 1, SILT_CONFINE_REFERENCE, SILT_PATH_EXACT. */
void silt_sheet_opts_default(silt_sheet_opts *opts);

/*
 * step_sheet (src/fluid_core.rs:770-785) on host slices of length len.
 * h, u, v are mutated in place; bed, nu, gate are read.  The planes are
 * copied to the GPU, stepped there and copied back.
 */
int silt_step_sheet(float *h, float *u, float *v, const float *bed, const float *nu, const uint8_t *gate,
                    size_t len, size_t res, float cell, float dt, silt_drive drive);

/* step_sheet_eps (src/fluid_core.rs:856-889) with every option. */
int silt_step_sheet_ex(float *h, float *u, float *v, const float *bed, const float *nu, const uint8_t *gate,
                       size_t len, size_t res, float cell, float dt, silt_drive drive,
                       const silt_sheet_opts *opts);

/* The same step on device pointers (hipMalloc'd, res * res each), enqueued on
 * hip_stream.  nu_uniform: if finite, every cell's viscosity equals it and nu
 * may be NULL (the water layer's case); NaN reads the nu plane.
 * The work-box plan reads the depth back to the host, so the call blocks
 * until the planes are final. */
int silt_step_sheet_device(float *h, float *u, float *v, const float *bed, const float *nu, float nu_uniform,
                           const uint8_t *gate, size_t res, float cell, float dt, silt_drive drive,
                           const silt_sheet_opts *opts, void *hip_stream);

/*
 * The walled sheet: step_sheet_eps_in (src/fluid_core.rs:856-889) with a wall
 * mask, solid = Some (the creek and ford scenes; FluidCore::step_pair passes
 * the mask of its hydro solids, :7954-7966).  solid: res * res bytes, nonzero
 * = wall.  Walls keep the wet-set work boxes and the substep loop on every
 * grid (no pace policy: opts.pace is ignored), Courant 0.42 for wind too, and
 * every substep adds the surface lift, the no-slip ghost, the redirect into
 * the flanks and the no-penetration filter (:1879-2082, :4225-4563).  The
 * stage kernels run it on every path (opts.path is ignored); opts.row_split
 * picks the reference's parallel-kick form on boxes of at least 48 x 32.
 * Returns the substep count.
 */
int silt_step_sheet_walled(float *h, float *u, float *v, const float *bed, const float *nu, const uint8_t *gate,
                           const uint8_t *solid, size_t len, size_t res, float cell, float dt, silt_drive drive,
                           const silt_sheet_opts *opts);

/* The walled step on device pointers (as silt_step_sheet_device; solid is a
 * device plane). */
int silt_step_sheet_walled_device(float *h, float *u, float *v, const float *bed, const float *nu, float nu_uniform,
                                  const uint8_t *gate, const uint8_t *solid, size_t res, float cell, float dt,
                                  silt_drive drive, const silt_sheet_opts *opts, void *hip_stream);

/* displace_occupied (src/fluid_core.rs:4382-4444) on host planes: water in
 * cells a body occupies is split over the open 4-neighbours (or the open
 * cells within two cells).  FluidCore::step_pair runs it on each layer before
 * the walled step. */
int silt_displace_occupied(float *h, float *u, float *v, const uint8_t *solid, size_t len, size_t res);

/* Library identity: "silt_amd <version> gfx950". */
const char *silt_version(void);

/* Rows of the grid each workgroup of the fused streaming kernel takes, for
 * sheets of side res; 0 restores the default (the smallest height that keeps
 * every tile of a launch resident at once).  Process-wide.  No counterpart in
 * the reference: a tuning knob of the GPU path (see silt_pair_autotune). */
int silt_set_chunk_rows(size_t res, int rows);
/* The rows pinned for side res, 0 when the default applies. */
int silt_chunk_rows(size_t res);

#ifdef __cplusplus
}
#endif

#endif /* SILT_SHEET_H */
