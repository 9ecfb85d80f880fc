/*
 * oracle/rut.c -- CPU restatement of the reference's Bekker rut coupling on
 * the sheet path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle/sheet.c).  Compiled twice, FP32 and
 * FP64.
 *
 * Reference: src/soil.rs of x3-team/Dirty-gen --
 *   bekker_load :1224-1245, sinkage_with :932-946, rut_depth_from_force
 *   :1036-1061, contact_length :1215-1218, terramech :825-840,
 *   softness_of :1100-1106; the loaded, lane, tread branch of stamp_wheel
 *   :2526-2879 that ThroughDrive::stamp_puddle_tires uses
 *   (src/world.rs:4846-4930), with excavate :1800-1809 and deposit_spoil
 *   :2881-2946 onto the berm band (no clods on that branch).
 * The grid is any square grid of `cell` metres with its corner at origin.
 * Cosmetic soil fields (kind, loose, darken, roughness) are not restated;
 * they do not reach the bed the sheet reads (bed + placed).
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
#define K(x) ((real)(float)(x))

#if defined(REAL_IS_DOUBLE)
#define R_SQRT sqrt
#define R_POW pow
#define R_FLOOR floor
#define R_FABS fabs
#define R_FMIN fmin
#define R_FMAX fmax
#define R_EXP exp
#else
#define R_SQRT sqrtf
#define R_POW powf
#define R_FLOOR floorf
#define R_FABS fabsf
#define R_FMIN fminf
#define R_FMAX fmaxf
#define R_EXP expf
#endif

static real clampr(real x, real lo, real hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

/* bekker_load (src/soil.rs:1224-1245): 8 midpoint samples of k_eq y^n over
 * the chord of a rigid wheel. */
real F(oracle_bekker_load)(real sinkage, real radius, real width, real k_eq, real n) {
    if (sinkage <= K(1.0e-5) || k_eq <= K(0.0) || width <= K(0.0)) return K(0.0);
    real z = R_FMIN(sinkage, radius * K(0.95));
    real half = R_SQRT(R_FMAX(K(2.0) * radius * z - z * z, K(0.0)));
    if (half < K(1.0e-5)) return K(0.0);
    const int samples = 8;
    real acc = K(0.0);
    for (int i = 0; i < samples; i++) {
        real t = ((real)i + K(0.5)) / (real)samples;
        real x = -half + t * (K(2.0) * half);
        real y = R_SQRT(R_FMAX(radius * radius - x * x, K(0.0))) - (radius - z);
        y = R_FMAX(y, K(0.0));
        acc += k_eq * R_POW(y, n);
    }
    real avg_p = acc / (real)samples;
    return avg_p * (K(2.0) * half) * width;
}

/* sinkage_with (src/soil.rs:932-946): 18 halvings of [0, 0.9 r]. */
real F(oracle_sinkage_with)(real load, real radius, real width, real k_eq, real n) {
    real lo = K(0.0), hi = radius * K(0.9);
    for (int k = 0; k < 18; k++) {
        real mid = K(0.5) * (lo + hi);
        real w = F(oracle_bekker_load)(mid, radius, width, k_eq, n);
        if (w < load)
            lo = mid;
        else
            hi = mid;
    }
    return K(0.5) * (lo + hi);
}

/* rut_depth_from_force (src/soil.rs:1036-1061) with the soil's modulus. */
real F(oracle_rut_depth_from_force)(real normal, real push, real radius, real width, real k_eq, real n) {
    real nn = R_FMAX(normal, K(0.0)), p = R_FMAX(push, K(0.0));
    real load = R_SQRT(nn * nn + p * p);
    if (load < K(1.0) || !isfinite(load)) return K(0.0);
    real r = R_FMAX(radius, K(0.05));
    real depth = F(oracle_sinkage_with)(load, r, R_FMAX(width, K(0.02)), k_eq, n);
    return R_FMIN(depth, r * K(0.85));
}

/* contact_length (src/soil.rs:1215-1218) */
real F(oracle_contact_length)(real radius, real sinkage) {
    real z = clampr(sinkage, K(0.0), radius * K(0.98));
    return K(2.0) * R_SQRT(R_FMAX(K(2.0) * radius * z - z * z, K(0.0)));
}

/* terramech(softness).k_eq and .n (src/soil.rs:825-840) */
void F(oracle_terramech)(real softness, real *k_eq, real *n) {
    real s = clampr(softness, K(0.0), K(1.0));
    *k_eq = K(320000.0) + (K(48000.0) - K(320000.0)) * s;
    *n = K(0.95) + (K(0.72) - K(0.95)) * s;
}

/* softness_of (src/soil.rs:1100-1106) */
real F(oracle_softness_of)(real wetness, real loose, real water) {
    real loose_n = clampr(loose / K(0.09), K(0.0), K(1.0));
    real depth = clampr(water / K(0.18), K(0.0), K(1.0));
    return clampr(clampr(wetness, K(0.0), K(1.0)) * K(0.48) + depth * K(0.34) + loose_n * K(0.22), K(0.0), K(1.0));
}

/* One contact: hub (cx, hub_y, cz), forward (fx, fz), right (rx, rz). */
typedef struct {
    real cx, cz, hub_y, fx, fz, rx, rz, radius, width, sinkage, normal, push, k_eq, n;
} Contact;

/*
 * stamp_wheel's loaded lane tread branch (src/soil.rs:2526-2879): the rut
 * floor is cut towards the Bekker depth of the contact force, placed soil
 * first (excavate), and the cut volume lands on the berm band beside the
 * tread (deposit_spoil).  Returns the cut volume, cubic metres.
 */
real F(oracle_stamp_tread)(real *bed, real *placed, const real *initial, int res, real cell, real ox, real oz,
                           const real *cv, real dt) {
    Contact c = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7], cv[8], cv[9], cv[10], cv[11], cv[12], cv[13]};
    int loaded = c.normal > K(1.0) || c.push > K(1.0);
    real forced = loaded ? F(oracle_rut_depth_from_force)(c.normal, c.push, c.radius, c.width, c.k_eq, c.n) : K(0.0);
    if (c.sinkage < K(0.004) && forced < K(0.004)) return K(0.0);
    real len = R_FMAX(F(oracle_contact_length)(c.radius, R_FMAX(c.sinkage, forced)), c.radius * K(0.35));
    real tread_core = R_FMAX(c.width * K(0.42), cell * K(0.45));
    real tread_edge = tread_core + cell * K(2.75);
    real half_w = tread_edge;
    real berm_band = cell * K(2.4);
    real half_l = R_FMAX(len * K(0.5) + cell * K(0.25), cell * K(1.05));
    real along_reach = half_l + cell * K(0.35);
    real lateral_reach = half_w + K(0.0) + berm_band + cell * K(0.4);
    real reach = along_reach + lateral_reach;
    long i0 = (long)R_FLOOR((c.cx - reach - ox) / cell), i1 = (long)R_FLOOR((c.cx + reach - ox) / cell);
    long j0 = (long)R_FLOOR((c.cz - reach - oz) / cell), j1 = (long)R_FLOOR((c.cz + reach - oz) / cell);
    long ilo = i0 > 1 ? i0 : 1, ihi = i1 < res - 1 ? i1 : res - 1;
    long jlo = j0 > 1 ? j0 : 1, jhi = j1 < res - 1 ? j1 : res - 1;
    real area = cell * cell, vol = K(0.0);
    int nb = 0, capb = 4096;
    int *berms = (int *)malloc(sizeof(int) * (size_t)capb);
    real want = loaded ? R_FMIN(forced, K(0.40)) : K(0.0);
    real fast = forced;
    for (long j = jlo; j < jhi; j++)
        for (long i = ilo; i < ihi; i++) {
            real x = ox + ((real)i + K(0.5)) * cell, z = oz + ((real)j + K(0.5)) * cell;
            real relx = x - c.cx, relz = z - c.cz;
            real lx = relx * c.rx + relz * c.rz;
            real lz = relx * c.fx + relz * c.fz;
            real u = clampr(R_FABS(lz) / R_FMAX(along_reach, K(0.05)), K(0.0), K(1.0));
            real along = u < K(0.78) ? K(1.0) : clampr((K(1.0) - u) / K(0.22), K(0.0), K(1.0));
            if (along <= K(0.0)) continue;
            int idx = (int)(j * res + i);
            real column = bed[idx] + placed[idx];
            real penetration = column - (c.hub_y - c.radius);
            if (penetration > c.sinkage + K(0.03) && R_FABS(lz) > cell * K(0.5)) continue;
            real ax = R_FABS(lx);
            real across;
            if (ax <= tread_core) {
                across = K(1.0);
            } else if (ax >= tread_edge) {
                across = K(0.0);
            } else {
                real t = (ax - tread_core) / R_FMAX(tread_edge - tread_core, K(1.0e-3));
                real s = t * t * (K(3.0) - K(2.0) * t);
                across = K(1.0) - s;
            }
            if (across > K(0.12)) {
                if (!loaded) continue;
                real already = R_FMAX(initial[idx] - bed[idx], K(0.0));
                real local_want = want * across;
                real local_groove = fast * across;
                if (already < local_want) {
                    real room = R_FMAX(R_FMIN(local_want - already, K(0.40) - already), K(0.0));
                    real rate = already < local_groove ? K(48.0) : K(2.2);
                    real drop = room * (K(1.0) - R_EXP(-dt * rate)) * along;
                    if (drop > K(0.0)) {
                        real from_placed = R_FMIN(drop, placed[idx]);
                        placed[idx] -= from_placed;
                        bed[idx] -= drop - from_placed;
                        vol += drop * area;
                    }
                }
            } else {
                real start = tread_core + cell * K(2.55);
                real end = start + cell * K(1.35);
                if (ax > start && ax <= end && along > K(0.25)) {
                    if (R_FABS(x) + K(0.06) < R_FABS(c.cx)) continue;
                    if (nb < capb) berms[nb++] = idx;
                }
            }
        }
    real cut = vol;
    if (vol > K(1.0e-8) && nb > 0) {
        /* deposit_spoil with no clods (src/soil.rs:2881-2946) */
        real left = vol, cap = K(0.08);
        int *open = (int *)malloc(sizeof(int) * (size_t)nb), *next = (int *)malloc(sizeof(int) * (size_t)nb);
        int no = nb;
        memcpy(open, berms, sizeof(int) * (size_t)nb);
        for (int round = 0; round < 6; round++) {
            if (left <= K(1.0e-8) || no == 0) break;
            real share = left / ((real)no * area);
            int nn = 0;
            for (int q = 0; q < no; q++) {
                int idx = open[q];
                real room = R_FMAX(cap - placed[idx], K(0.0));
                real add = R_FMIN(share, room);
                if (add > K(0.0)) {
                    placed[idx] += add;
                    left -= add * area;
                }
                if (placed[idx] < cap - K(1.0e-4)) next[nn++] = idx;
            }
            memcpy(open, next, sizeof(int) * (size_t)nn);
            no = nn;
        }
        if (left > K(1.0e-8)) {
            real add = left / ((real)nb * area);
            for (int q = 0; q < nb; q++) placed[berms[q]] += add;
        }
        free(open);
        free(next);
    }
    free(berms);
    return cut;
}
