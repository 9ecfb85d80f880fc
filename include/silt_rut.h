/*
 * silt_rut.h -- Bekker rut coupling of libsilt_amd.so.
 *
 * The tire's Bekker sinkage (src/soil.rs bekker_load :1224-1245,
 * sinkage_with :932-946, rut_depth_from_force :1036-1061) and the rut it cuts
 * into the ground under the sheet (stamp_wheel :2526-2879, the loaded lane
 * tread branch that ThroughDrive::stamp_puddle_tires drives,
 * src/world.rs:4846-4930).  The ground is bed + placed on a square grid; the
 * sheet reads it as its bed (pull_substrate, src/world.rs:5511-5537).
 * The soil modulus is one lumped k_eq (N/m^(n+2)) and exponent n: a
 * k_c, k_phi soil maps to k_eq = k_c / width + k_phi.
 */
#ifndef SILT_RUT_H
#define SILT_RUT_H

#include <stddef.h>
#include <stdint.h>

#include "silt_sheet.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct silt_tire_contact {
    float x, z;                 /* contact point, world metres */
    float hub_y;                /* hub height, world metres */
    float forward_x, forward_z; /* unit rolling direction */
    float right_x, right_z;     /* unit axle direction */
    float radius, width;        /* wheel, metres */
    float sinkage;              /* the solver's sinkage, metres */
    float normal, push;         /* soil reaction and horizontal push, N */
    float k_eq, n;              /* soil modulus and exponent */
} silt_tire_contact;

/* bekker_load (src/soil.rs:1224-1245), N. */
float silt_bekker_load(float sinkage, float radius, float width, float k_eq, float n);
/* rut_depth_from_force (src/soil.rs:1036-1061), metres. */
float silt_rut_depth(float normal, float push, float radius, float width, float k_eq, float n);

/* Stamps the contacts, in order, into device planes bed and placed
 * (res x res, cell metres, corner at origin); initial is the ground before
 * any rut (the reference's initial_height).  Asynchronous on hip_stream. */
int silt_rut_stamp_device(float *bed, float *placed, const float *initial, size_t res, float cell, float origin_x,
                          float origin_z, const silt_tire_contact *contacts, size_t n, float dt, void *hip_stream);

/* silt_rut_stamp_device without any memory allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * prepared contacts go to a fixed device array of SILT_NOALLOC_MAX_CONTACTS
 * records (88 KiB per GPU, held until the process ends) instead of a buffer
 * the library grows; more contacts return SILT_ETOOBIG.  Every call on one
 * GPU shares that array: calls that overlap on one GPU race (the call waits
 * on hip_stream for its own stamp before it returns, as
 * silt_rut_stamp_device does).  Same arguments and results as
 * silt_rut_stamp_device. */
#define SILT_NOALLOC_MAX_CONTACTS 1024
int silt_rut_stamp_device_noalloc(float *bed, float *placed, const float *initial, size_t res, float cell,
                                  float origin_x, float origin_z, const silt_tire_contact *contacts, size_t n,
                                  float dt, void *hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* SILT_RUT_H */
