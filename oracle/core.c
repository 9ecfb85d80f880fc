/*
 * oracle/core.c -- CPU restatement of the reference's two-layer FluidCore
 * step on the sheet path: the 40 m coarse window and the 2 cm patch under the
 * tires, water and slurry.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle/sheet.c).  Compiled twice, FP32 and
 * FP64, like sheet.c, whose step_sheet_ex it calls for every sheet.
 *
 * Reference: src/fluid_core.rs and src/fluid_core/near.rs of x3-team/Dirty-gen.
 * Restated: FluidCore::step (:6034-6128) on the sheet path -- refine
 * (:8358-8460), near field (near.rs:150-328), tire shoves (push_on
 * :8169-8239, apply_pushes :8152-8167), step_pair (:7942-8150) with hydro
 * solids, spray peel (:8241-8314), restrict_fine (:8462-8555), soften_border
 * (:9594-9620), scour_contacts (:7364-7421), the foam and stirred-silt
 * tracers (:6210-6750), the wheel-lifted suspended solids (:8750-9003); the
 * setters map_water_bed (:5329-5355), paint/pour (:5820-5855, :5424-5440),
 * lay_slurry_mask (:5857-5880), read_substrate_grid (:5928-5962); ledger()
 * (:6821-6883) for the two sheets, parked volume and spray; coarse_cell
 * (:6998-7034); recenter with parking (:5409-5422, :9004-9226); drain_over_bed
 * (:5123-5214); pond_ruts and seat_rut_water (:5509-5792).
 * Not restated (no sheet-path configuration enables them): wet bodies, waves,
 * open water, river reach, stream cycle, dust and fog.
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
#define R_EXP exp
#define R_CEIL ceil
#define R_SQRT sqrt
#define R_FLOOR floor
#define R_FABS fabs
#define R_FMIN fmin
#define R_FMAX fmax
#define R_HYPOT hypot
#define R_COPYSIGN copysign
#define R_SIN sin
#define R_COS cos
#define R_ACOS acos
#else
#define R_EXP expf
#define R_CEIL ceilf
#define R_SQRT sqrtf
#define R_FLOOR floorf
#define R_FABS fabsf
#define R_FMIN fminf
#define R_FMAX fmaxf
#define R_HYPOT hypotf
#define R_COPYSIGN copysignf
#define R_SIN sinf
#define R_COS cosf
#define R_ACOS acosf
#endif
#define GRAV K(9.81)
#define R_ISFINITE(x) isfinite(x)
/* Rust f32::clamp: NaN passes through. */
static inline real clampr(real x, real lo, real hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

#define WRES 200
#define FRES 160
#define WN (WRES * WRES)
#define FN (FRES * FRES)
#define COARSE K(0.20)
#define FINE K(0.02)
#define NEAR_SUB 10
#define NEAR_SUB2 100
#define MAX_PUSH 64

/* From sheet.c (same precision). */
int F(oracle_step_sheet_ex)(real *h, real *u, real *v, const real *bed, const real *nu, const uint8_t *gate,
                            size_t res, real cell, real dt, int drive_wind, real wind_x, real wind_z,
                            real vorticity, real eps, int banks, real sub_cap, int pace, int inner_par,
                            int confine_mode, size_t *boxes_out, size_t *nboxes_out);
real F(oracle_dynamic_viscosity)(int grain, real moisture);
real F(oracle_column_sum)(const real *h, size_t n, real cell);
int F(oracle_step_sheet_walled)(real *h, real *u, real *v, const real *bed, const real *nu, const uint8_t *gate,
                                const uint8_t *solid, size_t res, real cell, real dt, int drive_wind, real wind_x,
                                real wind_z, real vorticity, real eps, real sub_cap, int inner_par);
void F(oracle_displace_occupied)(real *h, real *u, real *v, const uint8_t *solid, size_t res, real *give);

/* HydroSolid (src/fluid_core.rs:219-238). */
typedef struct {
    real x, z, yaw, half_length, half_width, bottom_y, top_y;
    int contact;
    real radius, width;
} HSolid;
#define MAX_SOLIDS 64

typedef struct {
    real x, z, fx, fz, radius, width, speed, slip;
} Push;

/* One residual tile of the near field, keyed by the world 20 cm cell. */
typedef struct {
    int32_t kx, kz;
    int used;
    real w[NEAR_SUB2], s[NEAR_SUB2], uw[NEAR_SUB2], vw[NEAR_SUB2];
    int fresh;
    real idle_s, peak;
} Tile;

struct ParkedSlot {
    int32_t kx, kz;
    int state; /* 0 empty, 1 live, 2 deleted */
    real water, slurry, uw, vw, us, vs, moisture, foam, turbid, susp;
};

typedef struct {
    real origin[2];
    real *water, *slurry, *uw, *vw, *us, *vs, *bed, *moisture;
    uint8_t *grain;
    real fine_origin[2];
    real *fw, *fs, *fuw, *fvw, *fus, *fvs, *fbed, *fbed_s, *fmoist;
    uint8_t *fgrain;
    uint8_t *hold;
    int fine_on;
    int fine_pace;
    real fine_sub_cap;
    int coarse_confine; /* confine mode of the coarse sheets (0 reference = Gauss-Seidel here) */
    int near_on;
    Push pushes[MAX_PUSH];
    int npush;
    real spray_m3;
    real frozen_w, frozen_s; /* fine volume parked off the window, cubic metres */
    /* near field: open addressing over world keys */
    Tile *tiles;
    size_t tcap;
    size_t ntiles;
    /* scratch */
    real *nu_w, *nu_s;
    uint8_t *gate;
    int spray_count;
    real moisture_ambient; /* read_substrate's most common moisture */
    /* cells parked off the window by recenter: open addressing by world key */
    struct ParkedSlot *parked;
    size_t pcap, pcount, pused;
    /* walls (set_hydro_solids) and the scour account */
    HSolid solids[MAX_SOLIDS];
    int nsolids;
    real *bed_drop;
    real eroded_m3, scour_hole_m;
    uint8_t *mask, *fmask;
    real *give;
    /* tracers: foam fraction and stirred silt (m) on the window, the live
     * foam and this step's wet and shore cells, the silt accounts, the
     * pushes of the last step */
    real *foam, *turbid, *tr_s;
    uint8_t *foam_live, *wet_m, *shore_m;
    real stirred_m3, turbid_settled_m3;
    /* suspended solids lifted by the wheels (metres in the water column) and
     * their accounts: lifted from the bed, landed back */
    real *susp;
    real entrained_m3, settled_susp_m3;
    Push last_pushes[MAX_PUSH];
    int nlast;
} Core;

static real *zalloc(size_t n) { return (real *)calloc(n, sizeof(real)); }

Core *F(oracle_core_new)(real cx, real cz) {
    Core *c = (Core *)calloc(1, sizeof(Core));
    real extent = (real)WRES * COARSE;
    /* snap_origin (src/fluid_core.rs:9888-9892) */
    real x = cx - extent * K(0.5), z = cz - extent * K(0.5);
    c->origin[0] = R_FLOOR(x / COARSE) * COARSE;
    c->origin[1] = R_FLOOR(z / COARSE) * COARSE;
    c->water = zalloc(WN); c->slurry = zalloc(WN); c->uw = zalloc(WN); c->vw = zalloc(WN);
    c->us = zalloc(WN); c->vs = zalloc(WN); c->bed = zalloc(WN); c->moisture = zalloc(WN);
    c->grain = (uint8_t *)malloc(WN);
    memset(c->grain, 4, WN); /* Grain::Other */
    c->fw = zalloc(FN); c->fs = zalloc(FN); c->fuw = zalloc(FN); c->fvw = zalloc(FN);
    c->fus = zalloc(FN); c->fvs = zalloc(FN); c->fbed = zalloc(FN); c->fbed_s = zalloc(FN);
    c->fmoist = zalloc(FN);
    c->fgrain = (uint8_t *)malloc(FN);
    memset(c->fgrain, 4, FN);
    c->hold = (uint8_t *)calloc(WN, 1);
    c->fine_pace = 0;
    c->fine_sub_cap = K(8.0);
    c->near_on = 1;
    c->tcap = 1 << 14;
    c->tiles = (Tile *)calloc(c->tcap, sizeof(Tile));
    c->nu_w = zalloc(WN);
    c->nu_s = zalloc(WN);
    c->gate = (uint8_t *)malloc(WN);
    memset(c->gate, 1, WN);
    c->bed_drop = zalloc(WN);
    c->mask = (uint8_t *)calloc(WN, 1);
    c->fmask = (uint8_t *)calloc(FN, 1);
    c->give = zalloc(WN);
    c->foam = zalloc(WN);
    c->turbid = zalloc(WN);
    c->susp = zalloc(WN);
    c->tr_s = zalloc(2 * WN);
    c->foam_live = (uint8_t *)calloc(WN, 1);
    c->wet_m = (uint8_t *)calloc(WN, 1);
    c->shore_m = (uint8_t *)calloc(WN, 1);
    real nw = F(oracle_dynamic_viscosity)(3, K(1.0));
    for (size_t k = 0; k < WN; k++) c->nu_w[k] = nw;
    return c;
}

void F(oracle_core_free)(Core *c) {
    if (!c) return;
    real *planes[] = {c->water, c->slurry, c->uw, c->vw, c->us, c->vs, c->bed, c->moisture, c->fw, c->fs, c->fuw,
                      c->fvw, c->fus, c->fvs, c->fbed, c->fbed_s, c->fmoist, c->nu_w, c->nu_s};
    for (size_t k = 0; k < sizeof planes / sizeof planes[0]; k++) free(planes[k]);
    free(c->grain); free(c->fgrain); free(c->hold); free(c->tiles); free(c->gate); free(c->parked);
    free(c->bed_drop); free(c->mask); free(c->fmask); free(c->give);
    free(c->susp);
    free(c->foam); free(c->turbid); free(c->tr_s); free(c->foam_live); free(c->wet_m); free(c->shore_m);
    free(c);
}

void F(oracle_core_set_fine_pace)(Core *c, int pace) { c->fine_pace = pace > 3 ? 3 : pace; }
void F(oracle_core_set_fine_sub_cap)(Core *c, real cap) {
    c->fine_sub_cap = isfinite(cap) ? (cap < K(1.0) ? K(1.0) : (cap > K(8.0) ? K(8.0) : cap)) : K(8.0);
}
void F(oracle_core_set_coarse_confine)(Core *c, int mode) { c->coarse_confine = mode; }
void F(oracle_core_set_near_field)(Core *c, int on) {
    c->near_on = on;
    if (!on) {
        memset(c->tiles, 0, c->tcap * sizeof(Tile));
        c->ntiles = 0;
    }
}

/* ---- near field tile map --------------------------------------------------- */

static size_t tile_hash(int32_t kx, int32_t kz, size_t cap) {
    uint32_t h = (uint32_t)kx * 0x9E3779B1u ^ ((uint32_t)kz * 0x85EBCA6Bu + 0x27d4eb2fu);
    h ^= h >> 15;
    return (size_t)h & (cap - 1);
}

static Tile *tile_find(Core *c, int32_t kx, int32_t kz) {
    size_t i = tile_hash(kx, kz, c->tcap);
    for (size_t n = 0; n < c->tcap; n++, i = (i + 1) & (c->tcap - 1)) {
        Tile *t = &c->tiles[i];
        if (!t->used) return NULL;
        if (t->used == 1 && t->kx == kx && t->kz == kz) return t;
    }
    return NULL;
}

static void tile_remove(Core *c, int32_t kx, int32_t kz) {
    Tile *t = tile_find(c, kx, kz);
    if (t) {
        t->used = 2; /* tombstone */
        c->ntiles--;
    }
}

static Tile *tile_get_or_insert(Core *c, int32_t kx, int32_t kz) {
    Tile *t = tile_find(c, kx, kz);
    if (t) return t;
    size_t i = tile_hash(kx, kz, c->tcap);
    for (size_t n = 0; n < c->tcap; n++, i = (i + 1) & (c->tcap - 1)) {
        Tile *s = &c->tiles[i];
        if (s->used != 1) {
            memset(s, 0, sizeof *s);
            s->used = 1;
            s->kx = kx;
            s->kz = kz;
            c->ntiles++;
            return s;
        }
    }
    return NULL;
}

size_t F(oracle_core_near_tiles)(Core *c) { return c->ntiles; }

/* ---- geometry helpers ------------------------------------------------------ */

/* coarse_axis (src/fluid_core.rs:9557-9568); -1 outside the window. */
static int coarse_axis(real origin, real p) {
    real i = R_FLOOR((p - origin) / COARSE);
    if (i < K(0.0)) return -1;
    if (i >= (real)WRES) return -1;
    return (int)i;
}

static void coarse_xz(const Core *c, int i, int j, real *x, real *z) {
    *x = c->origin[0] + ((real)i + K(0.5)) * COARSE;
    *z = c->origin[1] + ((real)j + K(0.5)) * COARSE;
}

/* fine_coarse_axes (src/fluid_core.rs:9518-9534); -1 = off the window. */
static void fine_axes(const Core *c, int col[FRES], int row[FRES]) {
    for (int i = 0; i < FRES; i++)
        col[i] = coarse_axis(c->origin[0], c->fine_origin[0] + ((real)i + K(0.5)) * FINE);
    for (int j = 0; j < FRES; j++)
        row[j] = coarse_axis(c->origin[1], c->fine_origin[1] + ((real)j + K(0.5)) * FINE);
}

/* world_key (src/fluid_core.rs:9894-9896) */
static void world_key(real x, real z, int32_t *kx, int32_t *kz) {
    *kx = (int32_t)R_FLOOR(x / COARSE);
    *kz = (int32_t)R_FLOOR(z / COARSE);
}

/* index_span (src/fluid_core.rs:9877-9886) */
static void index_span(real origin, real cell, int res, real center, real reach, int *lo, int *hi) {
    if (!(cell > K(0.0)) || res == 0) {
        *lo = *hi = 0;
        return;
    }
    real l = R_FMAX(R_FLOOR((center - reach - origin) / cell - K(1.0)), K(0.0));
    real h = R_FMAX(R_CEIL((center + reach - origin) / cell + K(2.0)), K(0.0));
    long li = (long)l, hi_ = (long)h;
    if (li > res) li = res;
    if (hi_ > res) hi_ = res;
    if (hi_ < li) hi_ = li;
    *lo = (int)li;
    *hi = (int)hi_;
}

/* dist2_segment (src/fluid_core.rs:9648-9662) */
static real dist2_segment(real px, real pz, real ax, real az, real bx, real bz) {
    real abx = bx - ax, abz = bz - az;
    real ab2 = abx * abx + abz * abz;
    real t;
    if (ab2 < K(1.0e-8)) {
        t = K(0.0);
    } else {
        t = ((px - ax) * abx + (pz - az) * abz) / ab2;
        if (t < K(0.0)) t = K(0.0);
        if (t > K(1.0)) t = K(1.0);
    }
    real dx = px - (ax + abx * t), dz = pz - (az + abz * t);
    return dx * dx + dz * dz;
}

/* near_tire_zone (src/fluid_core.rs:9624-9646) */
static int near_tire_zone(const Push *p, int n, real x, real z, real r2) {
    for (int k = 0; k < n; k++) {
        real dx = x - p[k].x, dz = z - p[k].z;
        if (dx * dx + dz * dz <= r2) return 1;
    }
    for (int a = 0; a < n; a++)
        for (int b = a + 1; b < n; b++) {
            if (R_FABS(p[a].x - p[b].x) >= K(0.45)) continue;
            if (dist2_segment(x, z, p[a].x, p[a].z, p[b].x, p[b].z) <= r2) return 1;
        }
    return 0;
}

static int fine_held_at(const Core *c, real x, real z) {
    int i = coarse_axis(c->origin[0], x), j = coarse_axis(c->origin[1], z);
    if (i < 0 || j < 0) return 0;
    return c->hold[j * WRES + i] == 1;
}

/* cell_is_refined (src/fluid_core.rs:8337-8356) */
static int cell_is_refined(const Core *c, real x, real z) {
    if (!c->fine_on || !fine_held_at(c, x, z)) return 0;
    int i = coarse_axis(c->origin[0], x), j = coarse_axis(c->origin[1], z);
    real x0 = c->origin[0] + (real)i * COARSE, z0 = c->origin[1] + (real)j * COARSE;
    real x1 = x0 + COARSE, z1 = z0 + COARSE;
    real extent = (real)FRES * FINE;
    real r0 = c->fine_origin[0], r1 = c->fine_origin[0] + extent;
    real r2 = c->fine_origin[1], r3 = c->fine_origin[1] + extent;
    return x0 >= r0 - K(1.0e-4) && z0 >= r2 - K(1.0e-4) && x1 <= r1 + K(1.0e-4) && z1 <= r3 + K(1.0e-4);
}

/* ---- near field: src/fluid_core/near.rs:114-328 ------------------------------ */

typedef struct {
    int cidx;
    int32_t kx, kz;
    int i0, j0;
} NearCol;

/* near_columns (near.rs:152-183): held columns covered by exactly 10 x 10 fine cells. */
static int near_columns(const Core *c, NearCol *out) {
    int col[FRES], row[FRES];
    fine_axes(c, col, row);
    int cr[FRES][2], nc = 0, rr[FRES][2], nr = 0;
    for (int pass = 0; pass < 2; pass++) {
        int *ax = pass == 0 ? col : row;
        int(*dst)[2] = pass == 0 ? cr : rr;
        int *cnt = pass == 0 ? &nc : &nr;
        int i = 0;
        while (i < FRES) {
            int cc = ax[i], n = 1;
            while (i + n < FRES && ax[i + n] == cc) n++;
            if (cc != -1 && n == NEAR_SUB) {
                dst[*cnt][0] = cc;
                dst[*cnt][1] = i;
                (*cnt)++;
            }
            i += n;
        }
    }
    int n = 0;
    for (int a = 0; a < nr; a++)
        for (int b = 0; b < nc; b++) {
            int cidx = rr[a][0] * WRES + cr[b][0];
            if (c->hold[cidx] != 1) continue;
            real x, z;
            coarse_xz(c, cr[b][0], rr[a][0], &x, &z);
            out[n].cidx = cidx;
            world_key(x, z, &out[n].kx, &out[n].kz);
            out[n].i0 = cr[b][1];
            out[n].j0 = rr[a][1];
            n++;
        }
    return n;
}

static int near_inner(const Core *c, int cidx) {
    int i = cidx % WRES, j = cidx / WRES;
    if (i == 0 || j == 0 || i + 1 >= WRES || j + 1 >= WRES) return 0;
    return c->hold[cidx - 1] == 1 && c->hold[cidx + 1] == 1 && c->hold[cidx - WRES] == 1 &&
           c->hold[cidx + WRES] == 1;
}

/* recentre (near.rs:114-121): the mean is taken in f64. */
static void recentre(real *r) {
    double m = 0.0;
    for (int k = 0; k < NEAR_SUB2; k++) m += (double)r[k];
    real mean = (real)(m / (double)NEAR_SUB2);
    for (int k = 0; k < NEAR_SUB2; k++) r[k] -= mean;
}

/* abs_peak (near.rs:124-133) */
static real abs_peak(const real *r) {
    real hi[4] = {0, 0, 0, 0};
    for (int k = 0; k < NEAR_SUB2; k += 4)
        for (int l = 0; l < 4; l++) {
            real a = R_FABS(r[k + l]);
            hi[l] = a > hi[l] ? a : hi[l];
        }
    return R_FMAX(R_FMAX(R_FMAX(hi[0], hi[1]), hi[2]), hi[3]);
}

/* floor_scale (near.rs:136-148) */
static real floor_scale(real base, const real *r) {
    real low = K(0.0);
    for (int k = 0; k < NEAR_SUB2; k++) low = R_FMIN(low, r[k]);
    if (low >= K(0.0) || base + low >= K(0.0)) return K(1.0);
    if (base <= K(0.0)) return K(0.0);
    real s = base / -low;
    return s < K(0.0) ? K(0.0) : (s > K(1.0) ? K(1.0) : s);
}

static real tile_left(const Tile *t) {
    if (t->idle_s <= K(0.0)) return K(1.0);
    return R_EXP(-t->idle_s / K(1.5));
}

/* near_capture (near.rs:199-261) */
static void near_capture(Core *c) {
    if (!c->fine_on || !c->near_on) return;
    static NearCol cols[FN / NEAR_SUB2 + 4];
    int n = near_columns(c, cols);
    for (int q = 0; q < n; q++) {
        NearCol nc = cols[q];
        real sw = K(0.0), ss = K(0.0);
        for (int b = 0; b < NEAR_SUB; b++) {
            int f = (nc.j0 + b) * FRES + nc.i0;
            real rw = -K(0.0), rs = -K(0.0);
            for (int a = 0; a < NEAR_SUB; a++) {
                rw += c->fw[f + a];
                rs += c->fs[f + a];
            }
            sw += rw;
            ss += rs;
        }
        if (sw <= K(0.0) && ss <= K(0.0)) {
            tile_remove(c, nc.kx, nc.kz);
            continue;
        }
        real cnt = (real)NEAR_SUB2;
        real mw = sw / cnt, ms = ss / cnt;
        real rw[NEAR_SUB2], rs[NEAR_SUB2];
        for (int b = 0; b < NEAR_SUB; b++) {
            int f = (nc.j0 + b) * FRES + nc.i0;
            for (int a = 0; a < NEAR_SUB; a++) {
                rw[b * NEAR_SUB + a] = c->fw[f + a] - mw;
                rs[b * NEAR_SUB + a] = c->fs[f + a] - ms;
            }
        }
        real peak = R_FMAX(abs_peak(rw), abs_peak(rs));
        int inner = near_inner(c, nc.cidx);
        if (peak <= K(1.0e-5) && !inner) {
            tile_remove(c, nc.kx, nc.kz);
            continue;
        }
        recentre(rw);
        recentre(rs);
        Tile *t = tile_get_or_insert(c, nc.kx, nc.kz);
        if (!t) continue;
        memcpy(t->w, rw, sizeof rw);
        memcpy(t->s, rs, sizeof rs);
        if (inner && sw > K(1.0e-10)) {
            real mu = K(0.0), mv = K(0.0);
            for (int b = 0; b < NEAR_SUB; b++) {
                int f = (nc.j0 + b) * FRES + nc.i0;
                for (int a = 0; a < NEAR_SUB; a++) {
                    real w = c->fw[f + a];
                    mu += w * c->fuw[f + a];
                    mv += w * c->fvw[f + a];
                }
            }
            real uw = mu / sw, vw = mv / sw;
            for (int b = 0; b < NEAR_SUB; b++) {
                int f = (nc.j0 + b) * FRES + nc.i0;
                for (int a = 0; a < NEAR_SUB; a++) {
                    int wet = c->fw[f + a] > K(1.0e-6);
                    t->uw[b * NEAR_SUB + a] = wet ? c->fuw[f + a] - uw : K(0.0);
                    t->vw[b * NEAR_SUB + a] = wet ? c->fvw[f + a] - vw : K(0.0);
                }
            }
        } else {
            memset(t->uw, 0, sizeof t->uw);
            memset(t->vw, 0, sizeof t->vw);
        }
        t->peak = peak;
        t->idle_s = K(0.0);
        t->fresh = 1;
    }
}

/* near_fade (near.rs:300-313) */
static void near_fade(Core *c, real dt) {
    if (!c->ntiles) return;
    if (!(dt > K(0.0) && isfinite(dt))) dt = K(0.0);
    for (size_t k = 0; k < c->tcap; k++) {
        Tile *t = &c->tiles[k];
        if (t->used != 1) continue;
        if (t->fresh) {
            t->fresh = 0;
            continue;
        }
        t->idle_s += dt;
        if (!(t->peak * tile_left(t) > K(1.0e-5))) {
            t->used = 2;
            c->ntiles--;
        }
    }
}

/* near_apply (near.rs:265-297) */
static void near_apply(Core *c, real dt) {
    if (c->ntiles) {
        static NearCol cols[FN / NEAR_SUB2 + 4];
        int n = near_columns(c, cols);
        for (int q = 0; q < n; q++) {
            NearCol nc = cols[q];
            Tile *t = tile_find(c, nc.kx, nc.kz);
            if (!t) continue;
            int f0 = nc.j0 * FRES + nc.i0;
            real bw = c->fw[f0], bs = c->fs[f0];
            int inner = near_inner(c, nc.cidx);
            real left = tile_left(t);
            real aw = left * floor_scale(bw / R_FMAX(left, K(1.0e-6)), t->w);
            real as = left * floor_scale(bs / R_FMAX(left, K(1.0e-6)), t->s);
            for (int b = 0; b < NEAR_SUB; b++)
                for (int a = 0; a < NEAR_SUB; a++) {
                    int f = (nc.j0 + b) * FRES + nc.i0 + a;
                    int k = b * NEAR_SUB + a;
                    real w = R_FMAX(bw + aw * t->w[k], K(0.0));
                    real s = R_FMAX(bs + as * t->s[k], K(0.0));
                    c->fw[f] = w;
                    c->fs[f] = s;
                    if (inner && w > K(1.0e-6)) {
                        c->fuw[f] += left * t->uw[k];
                        c->fvw[f] += left * t->vw[k];
                    }
                }
        }
    }
    near_fade(c, dt);
}

/* ---- refine / restrict: src/fluid_core.rs:8358-8555 -------------------------- */

static void restrict_fine(Core *c) {
    if (!c->fine_on) return;
    near_capture(c);
    const real fine_area = FINE * FINE, coarse_area = COARSE * COARSE;
    int col[FRES], row[FRES];
    fine_axes(c, col, row);
    static real aw[WN], as[WN], auw[WN], avw[WN], aus[WN], avs[WN];
    static uint32_t cnt[WN];
    static int touch[WN];
    int ntouch = 0;
    for (int j = 0; j < FRES; j++) {
        int cj = row[j];
        for (int i = 0; i < FRES; i++) {
            int ci = col[i];
            int fidx = j * FRES + i;
            if (cj == -1 || ci == -1) {
                /* Off the coarse window: park the fine column's volume. */
                c->frozen_w += c->fw[fidx] * fine_area / coarse_area * coarse_area;
                c->frozen_s += c->fs[fidx] * fine_area / coarse_area * coarse_area;
                continue;
            }
            int cidx = cj * WRES + ci;
            if (cnt[cidx] == 0) touch[ntouch++] = cidx;
            cnt[cidx]++;
            real fw = c->fw[fidx], fs = c->fs[fidx];
            if (fw == K(0.0) && fs == K(0.0)) continue;
            aw[cidx] += c->fw[fidx] * fine_area;
            as[cidx] += c->fs[fidx] * fine_area;
            auw[cidx] += c->fuw[fidx] * c->fw[fidx] * fine_area;
            avw[cidx] += c->fvw[fidx] * c->fw[fidx] * fine_area;
            aus[cidx] += c->fus[fidx] * c->fs[fidx] * fine_area;
            avs[cidx] += c->fvs[fidx] * c->fs[fidx] * fine_area;
        }
    }
    const uint32_t full = 100;
    for (int t = 0; t < ntouch; t++) {
        int idx = touch[t];
        int held = c->hold[idx] == 1;
        real w = aw[idx] / coarse_area, s = as[idx] / coarse_area;
        if (!held && cnt[idx] >= full) {
            if (!(w <= K(0.0) && s <= K(0.0))) {
                real old_w = c->water[idx], old_s = c->slurry[idx];
                real new_w = old_w + w, new_s = old_s + s;
                if (new_w > K(1.0e-6)) {
                    c->uw[idx] = (old_w * c->uw[idx] * coarse_area + auw[idx]) / (new_w * coarse_area);
                    c->vw[idx] = (old_w * c->vw[idx] * coarse_area + avw[idx]) / (new_w * coarse_area);
                }
                if (new_s > K(1.0e-6)) {
                    c->us[idx] = (old_s * c->us[idx] * coarse_area + aus[idx]) / (new_s * coarse_area);
                    c->vs[idx] = (old_s * c->vs[idx] * coarse_area + avs[idx]) / (new_s * coarse_area);
                }
                c->water[idx] = new_w;
                c->slurry[idx] = new_s;
            }
        } else {
            c->water[idx] = w;
            c->slurry[idx] = s;
            if (w > K(1.0e-6)) {
                c->uw[idx] = auw[idx] / (w * coarse_area);
                c->vw[idx] = avw[idx] / (w * coarse_area);
            }
            if (s > K(1.0e-6)) {
                c->us[idx] = aus[idx] / (s * coarse_area);
                c->vs[idx] = avs[idx] / (s * coarse_area);
            }
        }
        aw[idx] = as[idx] = auw[idx] = avw[idx] = aus[idx] = avs[idx] = K(0.0);
        cnt[idx] = 0;
    }
    memset(c->fw, 0, FN * sizeof(real));
    memset(c->fs, 0, FN * sizeof(real));
    c->fine_on = 0;
}

static void refine(Core *c, const Push *p, int np, real dt) {
    restrict_fine(c);
    if (np == 0) {
        near_fade(c, dt);
        return;
    }
    real cx = K(0.0), cz = K(0.0);
    for (int k = 0; k < np; k++) {
        cx += p[k].x;
        cz += p[k].z;
    }
    cx /= (real)np;
    cz /= (real)np;
    real extent = (real)FRES * FINE;
    c->fine_origin[0] = R_FLOOR((cx - extent * K(0.5)) / COARSE) * COARSE;
    c->fine_origin[1] = R_FLOOR((cz - extent * K(0.5)) / COARSE) * COARSE;
    memset(c->fw, 0, FN * sizeof(real));
    memset(c->fs, 0, FN * sizeof(real));
    memset(c->fuw, 0, FN * sizeof(real));
    memset(c->fvw, 0, FN * sizeof(real));
    memset(c->fus, 0, FN * sizeof(real));
    memset(c->fvs, 0, FN * sizeof(real));
    memset(c->hold, 0, WN);
    const uint32_t full = 100;
    real r2 = K(0.56) * K(0.56);
    int col[FRES], row[FRES];
    fine_axes(c, col, row);
    static uint32_t cnt[WN];
    static int touch[WN];
    int ntouch = 0;
    for (int j = 0; j < FRES; j++) {
        if (row[j] == -1) continue;
        for (int i = 0; i < FRES; i++) {
            if (col[i] == -1) continue;
            int cidx = row[j] * WRES + col[i];
            if (cnt[cidx] == 0) touch[ntouch++] = cidx;
            cnt[cidx]++;
        }
    }
    for (int t = 0; t < ntouch; t++) {
        int cidx = touch[t];
        if (cnt[cidx] != full) continue;
        real x, z;
        coarse_xz(c, cidx % WRES, cidx / WRES, &x, &z);
        if (near_tire_zone(p, np, x, z, r2)) c->hold[cidx] = 1;
    }
    for (int j = 0; j < FRES; j++) {
        if (row[j] == -1) continue;
        for (int i = 0; i < FRES; i++) {
            if (col[i] == -1) continue;
            int fidx = j * FRES + i, cidx = row[j] * WRES + col[i];
            c->fmoist[fidx] = c->moisture[cidx];
            c->fgrain[fidx] = c->grain[cidx];
            if (c->hold[cidx] != 1) {
                c->fbed[fidx] = c->bed[cidx] + c->water[cidx];
                c->fbed_s[fidx] = c->bed[cidx] + c->slurry[cidx];
                continue;
            }
            c->fbed[fidx] = c->bed[cidx];
            c->fbed_s[fidx] = c->bed[cidx];
            c->fw[fidx] = c->water[cidx];
            c->fs[fidx] = c->slurry[cidx];
            c->fuw[fidx] = c->uw[cidx];
            c->fvw[fidx] = c->vw[cidx];
            c->fus[fidx] = c->us[cidx];
            c->fvs[fidx] = c->vs[cidx];
        }
    }
    for (int t = 0; t < ntouch; t++) {
        int cidx = touch[t];
        if (c->hold[cidx] == 1)
            c->water[cidx] = c->slurry[cidx] = c->uw[cidx] = c->vw[cidx] = c->us[cidx] = c->vs[cidx] = K(0.0);
        cnt[cidx] = 0;
    }
    near_apply(c, dt);
    c->fine_on = 1;
}

/* ---- tire shoves and spray -------------------------------------------------- */

/* push_on (src/fluid_core.rs:8169-8239) */
static void push_on(Core *c, const Push *p, int fine) {
    int res = fine ? FRES : WRES;
    real cell = fine ? FINE : COARSE;
    const real *origin = fine ? c->fine_origin : c->origin;
    real *water = fine ? c->fw : c->water, *slurry = fine ? c->fs : c->slurry;
    real *uw = fine ? c->fuw : c->uw, *vw = fine ? c->fvw : c->vw;
    real *us = fine ? c->fus : c->us, *vs = fine ? c->fvs : c->vs;
    real len = R_FMAX(R_SQRT(p->fx * p->fx + p->fz * p->fz), K(1.0e-4));
    real fx = p->fx / len, fz = p->fz / len;
    real rx = fz, rz = -fx;
    real speed = R_FMAX(p->speed, K(0.0));
    real reach = R_HYPOT(p->radius * K(1.4), p->width * K(0.65) + cell) + cell;
    int i0, i1, j0, j1;
    index_span(origin[0], cell, res, p->x, reach, &i0, &i1);
    index_span(origin[1], cell, res, p->z, reach, &j0, &j1);
    for (int j = j0; j < j1; j++)
        for (int i = i0; i < i1; i++) {
            real x = origin[0] + ((real)i + K(0.5)) * cell;
            real z = origin[1] + ((real)j + K(0.5)) * cell;
            real dx = x - p->x, dz = z - p->z;
            real along = dx * fx + dz * fz;
            real lat = dx * rx + dz * rz;
            if (R_FABS(lat) > p->width * K(0.65) + cell) continue;
            if (along < -p->radius * K(1.4) || along > p->radius * K(1.1)) continue;
            int idx = j * res + i;
            real fl = R_FABS(lat) / (p->width * K(0.65) + cell);
            real flank = fl < K(0.0) ? K(0.0) : (fl > K(1.0) ? K(1.0) : fl);
            real shove = R_FMAX(speed, K(0.4)) * (K(1.0) - flank * flank);
            real sg = R_COPYSIGN(K(1.0), lat);
            if (water[idx] > K(0.004)) {
                uw[idx] += fx * shove * K(0.85);
                vw[idx] += fz * shove * K(0.85);
                uw[idx] += rx * sg * shove * K(0.18) * flank;
                vw[idx] += rz * sg * shove * K(0.18) * flank;
            }
            if (slurry[idx] > K(0.004)) {
                real fore = along > K(0.0) ? K(1.0) : K(0.72);
                us[idx] += fx * shove * fore;
                vs[idx] += fz * shove * fore;
                us[idx] += rx * sg * shove * K(0.28) * flank;
                vs[idx] += rz * sg * shove * K(0.28) * flank;
            }
        }
}

/* coarse_held (src/fluid_core.rs:9664-9672) */
static int coarse_held(const Core *c, real x, real z) {
    real i = R_FLOOR((x - c->origin[0]) / COARSE), j = R_FLOOR((z - c->origin[1]) / COARSE);
    if (i < K(0.0) || j < K(0.0)) return 0;
    if (i >= (real)WRES || j >= (real)WRES) return 0;
    return c->hold[(int)j * WRES + (int)i] == 1;
}

/* peel_spray (src/fluid_core.rs:8241-8314); returns the volume peeled. */
static real peel_spray(Core *c, int fine, int skip_held, int *events) {
    int res = fine ? FRES : WRES;
    real cell = fine ? FINE : COARSE;
    const real *origin = fine ? c->fine_origin : c->origin;
    real *water = fine ? c->fw : c->water;
    real area = cell * cell, total = K(0.0);
    for (int q = 0; q < c->npush; q++) {
        const Push *p = &c->pushes[q];
        if (p->speed < K(0.8)) continue;
        real reach = p->radius * K(1.3) + cell;
        int i0, i1, j0, j1;
        index_span(origin[0], cell, res, p->x, reach, &i0, &i1);
        index_span(origin[1], cell, res, p->z, reach, &j0, &j1);
        for (int j = j0; j < j1; j++)
            for (int i = i0; i < i1; i++) {
                real x = origin[0] + ((real)i + K(0.5)) * cell;
                real z = origin[1] + ((real)j + K(0.5)) * cell;
                if (skip_held && coarse_held(c, x, z)) continue;
                real dx = x - p->x, dz = z - p->z;
                if (dx * dx + dz * dz > (p->radius * K(1.3)) * (p->radius * K(1.3))) continue;
                int idx = j * res + i;
                if (water[idx] < K(0.01)) continue;
                real take = R_FMIN(water[idx] * K(0.04), K(0.004));
                if (take * area < K(1.0e-6)) continue;
                water[idx] -= take;
                total += take * area;
                c->spray_m3 += take * area;
                (*events)++;
            }
    }
    return total;
}

/* soften_border (src/fluid_core.rs:9594-9620), no open reach. */
static void soften_border(real *u, real *v, int res) {
    int edge = res / 12 > 2 ? res / 12 : 2;
    for (int j = 0; j < res; j++)
        for (int i = 0; i < res; i++) {
            int di = i < res - 1 - i ? i : res - 1 - i;
            int dj = j < res - 1 - j ? j : res - 1 - j;
            int d = di < dj ? di : dj;
            if (d >= edge) continue;
            real t = (real)d / (real)edge;
            real w = t * t * (K(3.0) - K(2.0) * t);
            u[j * res + i] *= w;
            v[j * res + i] *= w;
        }
}

/* wet_bounds (src/fluid_core.rs:794-819) then the slurry viscosity fill. */
static void fill_nu_s(Core *c, const real *slurry, const real *moisture, const uint8_t *grain, int res) {
    int i0 = res, i1 = 0, j0 = res, j1 = 0, any = 0;
    for (int j = 0; j < res; j++)
        for (int i = 0; i < res; i++)
            if (slurry[j * res + i] > K(0.0)) {
                any = 1;
                if (i < i0) i0 = i;
                if (i > i1) i1 = i;
                if (j < j0) j0 = j;
                if (j > j1) j1 = j;
            }
    if (!any) return;
    i0 = i0 >= 8 ? i0 - 8 : 0;
    j0 = j0 >= 8 ? j0 - 8 : 0;
    i1 = i1 + 8 < res - 1 ? i1 + 8 : res - 1;
    j1 = j1 + 8 < res - 1 ? j1 + 8 : res - 1;
    for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++)
            c->nu_s[j * res + i] = F(oracle_dynamic_viscosity)(grain[j * res + i], moisture[j * res + i]);
}


/* ---- walls: src/fluid_core.rs:4584-4616, 7164-7421, 7942-8110 ------------ */

/* solid_contains, :4584-4594 */
static int solid_contains(const HSolid *s, real x, real z) {
    real dx = x - s->x, dz = z - s->z;
    real fwd_x = R_SIN(s->yaw), fwd_z = R_COS(s->yaw);
    real right_x = fwd_z, right_z = -fwd_x;
    real along = dx * fwd_x + dz * fwd_z;
    real lat = dx * right_x + dz * right_z;
    return R_FABS(along) <= R_FMAX(s->half_length, K(0.0)) && R_FABS(lat) <= R_FMAX(s->half_width, K(0.0));
}

/* solid_overlaps, :4596-4604 */
static int solid_overlaps(const HSolid *s, real x, real z, real eta, real bed) {
    if (s->half_length <= K(1.0e-4) || s->half_width <= K(1.0e-4)) return 0;
    if (!solid_contains(s, x, z)) return 0;
    return eta > s->bottom_y + K(1.0e-4) && bed < s->top_y;
}

/* solid_mask, :7288-7362: walls 2 first, then contacts 1 (a contact wins a
 * shared cell); a cell is a wall while the free surface beside it stands
 * above the solid's underside. */
static void solid_mask(const Core *c, int fine, uint8_t *mask) {
    int res = fine ? FRES : WRES;
    real cell = fine ? FINE : COARSE;
    const real *origin = fine ? c->fine_origin : c->origin;
    const real *water = fine ? c->fw : c->water;
    const real *bed = fine ? c->fbed : c->bed;
    memset(mask, 0, (size_t)res * res);
    for (int pass = 0; pass < 2; pass++) {
        int contact = pass == 1;
        uint8_t value = contact ? 1 : 2;
        for (int k = 0; k < c->nsolids; k++) {
            const HSolid *s = &c->solids[k];
            if ((s->contact != 0) != contact) continue;
            if (s->half_length <= K(1.0e-4) || s->half_width <= K(1.0e-4)) continue;
            real reach = R_HYPOT(R_FMAX(s->half_length, K(0.0)), R_FMAX(s->half_width, K(0.0))) + cell;
            if (!R_ISFINITE(reach) || !R_ISFINITE(s->x) || !R_ISFINITE(s->z)) continue;
            real lo_i = R_FLOOR((s->x - reach - origin[0]) / cell - K(0.5));
            real hi_i = R_CEIL((s->x + reach - origin[0]) / cell - K(0.5));
            real lo_j = R_FLOOR((s->z - reach - origin[1]) / cell - K(0.5));
            real hi_j = R_CEIL((s->z + reach - origin[1]) / cell - K(0.5));
            real top = (real)(res - 1);
            if (hi_i < K(0.0) || hi_j < K(0.0) || lo_i > top || lo_j > top) continue;
            int i0 = (int)R_FMAX(lo_i, K(0.0)), i1 = (int)R_FMIN(hi_i, top);
            int j0 = (int)R_FMAX(lo_j, K(0.0)), j1 = (int)R_FMIN(hi_j, top);
            for (int j = j0; j <= j1; j++)
                for (int i = i0; i <= i1; i++) {
                    size_t idx = (size_t)j * res + i;
                    if (mask[idx] == value) continue;
                    real x = origin[0] + ((real)i + K(0.5)) * cell;
                    real z = origin[1] + ((real)j + K(0.5)) * cell;
                    real eta = bed[idx] + water[idx];
                    static const int nb[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
                    for (int q = 0; q < 4; q++) {
                        int ni = i + nb[q][0], nj = j + nb[q][1];
                        if (ni < 0 || nj < 0 || ni >= res || nj >= res) continue;
                        size_t nidx = (size_t)nj * res + ni;
                        eta = R_FMAX(eta, bed[nidx] + water[nidx]);
                    }
                    if (solid_overlaps(s, x, z, eta, bed[idx])) mask[idx] = value;
                }
        }
    }
}

/* One layer of step_pair with walls (:8056-8125): soften (coarse), the body
 * displacement, then the walled step. */
static void step_layer_walled(Core *c, int fine, int slurry, const uint8_t *mask, real dt) {
    int res = fine ? FRES : WRES;
    real cell = fine ? FINE : COARSE;
    real cap = fine ? c->fine_sub_cap : K(8.0);
    real *h = slurry ? (fine ? c->fs : c->slurry) : (fine ? c->fw : c->water);
    real *u = slurry ? (fine ? c->fus : c->us) : (fine ? c->fuw : c->uw);
    real *v = slurry ? (fine ? c->fvs : c->vs) : (fine ? c->fvw : c->vw);
    real *bed = slurry ? (fine ? c->fbed_s : c->bed) : (fine ? c->fbed : c->bed);
    if (!fine) soften_border(u, v, res);
    F(oracle_displace_occupied)(h, u, v, mask, (size_t)res, c->give);
    F(oracle_step_sheet_walled)(h, u, v, bed, slurry ? c->nu_s : c->nu_w, c->gate, mask, (size_t)res, cell, dt, 0,
                                K(0.0), K(0.0), slurry ? K(0.05) : K(0.08), K(1.0e-6), cap, 0);
}

/* scour_contacts, :7364-7421 */
static void scour_contacts(Core *c, real dt) {
    if (dt <= K(0.0) || c->nsolids == 0) return;
    solid_mask(c, 0, c->mask);
    const uint8_t *mask = c->mask;
    real area = COARSE * COARSE;
    real *cut = c->give;
    memset(cut, 0, WN * sizeof(real));
    for (int j = 1; j < WRES - 1; j++)
        for (int i = 1; i < WRES - 1; i++) {
            size_t idx = (size_t)j * WRES + i;
            if (mask[idx] != 0 || c->water[idx] < K(0.02)) continue;
            if (!(mask[idx - 1] == 1 || mask[idx + 1] == 1 || mask[idx - WRES] == 1 || mask[idx + WRES] == 1))
                continue;
            real speed = R_HYPOT(c->uw[idx], c->vw[idx]);
            if (!R_ISFINITE(speed)) continue;
            real tau = K(1000.0) * K(0.03) * speed * speed;
            real tau_c = K(1.5);
            if (tau <= tau_c) continue;
            real rate = R_FMIN(K(0.0035) * (tau / tau_c - K(1.0)), K(0.012));
            real room = R_FMAX(K(0.08) - c->bed_drop[idx], K(0.0));
            cut[idx] = R_FMIN(rate * dt, room);
        }
    for (size_t idx = 0; idx < WN; idx++) {
        real dh = cut[idx];
        if (dh <= K(0.0)) continue;
        c->bed[idx] -= dh;
        c->slurry[idx] += dh;
        c->us[idx] = c->uw[idx];
        c->vs[idx] = c->vw[idx];
        c->grain[idx] = 1; /* Grain::Silt */
        c->moisture[idx] = R_FMAX(c->moisture[idx], K(0.9));
        c->bed_drop[idx] += dh;
        c->eroded_m3 += dh * area;
        c->scour_hole_m = R_FMAX(c->scour_hole_m, c->bed_drop[idx]);
    }
}

/* cylinder_segment_volume, :4606-4615 */
static real cylinder_segment_volume(real depth, real radius, real width) {
    real d = clampr(depth, K(0.0), radius * K(2.0));
    if (d <= K(1.0e-4) || width <= K(0.0) || radius <= K(1.0e-3)) return K(0.0);
    real cc = clampr((radius - d) / radius, K(-1.0), K(1.0));
    real a = radius * radius * R_ACOS(cc) - (radius - d) * R_SQRT(R_FMAX(K(2.0) * radius * d - d * d, K(0.0)));
    return R_FMAX(a, K(0.0)) * width;
}

/* ring_surface, :7251-7286: mean bed and water of the wet cells around a
 * solid (outside it, within its reach). */
static int ring_surface(const Core *c, const HSolid *s, real *bed_out, real *water_out) {
    real reach = R_FMAX(s->half_length, s->half_width) + K(0.55);
    int i0 = (int)R_FLOOR((s->x - reach - c->origin[0]) / COARSE);
    int i1 = (int)R_FLOOR((s->x + reach - c->origin[0]) / COARSE);
    int j0 = (int)R_FLOOR((s->z - reach - c->origin[1]) / COARSE);
    int j1 = (int)R_FLOOR((s->z + reach - c->origin[1]) / COARSE);
    real bed = K(0.0), water = K(0.0), n = K(0.0);
    for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) {
            if (i < 0 || j < 0 || i >= WRES || j >= WRES) continue;
            real x, z;
            coarse_xz(c, i, j, &x, &z);
            if (solid_contains(s, x, z)) continue;
            if (R_HYPOT(x - s->x, z - s->z) > reach) continue;
            size_t idx = (size_t)j * WRES + i;
            if (c->water[idx] < K(0.01)) continue;
            bed += c->bed[idx];
            water += c->water[idx];
            n += K(1.0);
        }
    if (n < K(1.0)) return 0;
    *bed_out = bed / n;
    *water_out = water / n;
    return 1;
}

/* ---- tracers: foam and stirred silt on the window, src/fluid_core.rs:
 * 6210-6750, 9676-9880 --------------------------------------------------- */

#define FOAM_LIFE_S K(1.8)
#define FOAM_EQ K(0.85)
#define FOAM_TAU_S K(0.35)
#define SILT_SETTLE_S K(4.0)
#define SILT_COLUMN_CAP_M K(0.028)

/* smoothstep (:9688-9691) */
static real smooth_step(real e0, real e1, real x) {
    real t = clampr((x - e0) / (e1 - e0), K(0.0), K(1.0));
    return t * t * (K(3.0) - K(2.0) * t);
}

/* breaking_foam (:9676-9686): Froude foam of a film. */
static real breaking_foam(real speed, real thickness) {
    if (!R_ISFINITE(speed) || !R_ISFINITE(thickness) || thickness < K(0.004) || speed <= K(0.05)) return K(0.0);
    real froude = speed / R_SQRT(GRAV * thickness);
    return smooth_step(K(0.85), K(1.45), froude);
}

/* sample_foam (:9693-9714): bilinear in index space. */
static real sample_foam(const real *src, int res, real x, real z) {
    if (!R_ISFINITE(x) || !R_ISFINITE(z) || x < K(0.0) || z < K(0.0)) return K(0.0);
    size_t i0 = (size_t)R_FLOOR(x), j0 = (size_t)R_FLOOR(z);
    if (i0 >= (size_t)res || j0 >= (size_t)res) return K(0.0);
    size_t i1 = i0 + 1 < (size_t)res ? i0 + 1 : (size_t)res - 1;
    size_t j1 = j0 + 1 < (size_t)res ? j0 + 1 : (size_t)res - 1;
    real tx = clampr(x - (real)i0, K(0.0), K(1.0)), tz = clampr(z - (real)j0, K(0.0), K(1.0));
    real a = src[j0 * res + i0], b = src[j0 * res + i1], cc = src[j1 * res + i0], d = src[j1 * res + i1];
    real v = a + (b - a) * tx + (cc - a) * tz + (a - b - cc + d) * tx * tz;
    return R_ISFINITE(v) ? clampr(v, K(0.0), K(1.0)) : K(0.0);
}

/* turbid_span (:9766-9794): cells holding silt, grown by one. */
static int turbid_span(const real *t, int res, int *i0, int *i1, int *j0, int *j1) {
    int a = res, b = 0, cmin = res, d = 0, any = 0;
    for (int j = 0; j < res; j++) {
        int first = -1, last = -1;
        for (int i = 0; i < res; i++)
            if (t[j * res + i] > K(1.0e-8)) {
                if (first < 0) first = i;
                last = i;
            }
        if (first < 0) continue;
        any = 1;
        if (first < a) a = first;
        if (last > b) b = last;
        if (j < cmin) cmin = j;
        if (j > d) d = j;
    }
    if (!any) return 0;
    *i0 = a > 0 ? a - 1 : 0;
    *i1 = b + 1 < res - 1 ? b + 1 : res - 1;
    *j0 = cmin > 0 ? cmin - 1 : 0;
    *j1 = d + 1 < res - 1 ? d + 1 : res - 1;
    return 1;
}

/* shore_wet_cell (:9752-9764) */
static int shore_wet_cell(const real *water, int res, int i, int j) {
    if (i == 0 || j == 0 || i + 1 >= res || j + 1 >= res) return 0;
    static const int nb[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    for (int q = 0; q < 4; q++) {
        real nh = water[(j + nb[q][1]) * res + i + nb[q][0]];
        if (R_ISFINITE(nh) && nh < K(0.004)) return 1;
    }
    return 0;
}

/* shore_cell (:9826-9874): foam production and silt lift rate at a bank. */
static void shore_cell(const real *water, const real *uw, const real *vw, int res, int idx, real *foam_out,
                       real *rate_out) {
    *foam_out = K(0.0);
    *rate_out = K(0.0);
    int i = idx % res, j = idx / res;
    if (i == 0 || j == 0 || i + 1 >= res || j + 1 >= res) return;
    real h = water[idx];
    if (!R_ISFINITE(h) || h < K(0.004)) return;
    real nx = K(0.0), nz = K(0.0);
    int edge = 0;
    static const int nb[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    for (int q = 0; q < 4; q++) {
        real nh = water[(j + nb[q][1]) * res + i + nb[q][0]];
        if (R_ISFINITE(nh) && nh < K(0.004)) {
            edge = 1;
            nx += (real)nb[q][0];
            nz += (real)nb[q][1];
        }
    }
    if (!edge) return;
    real u = uw[idx], v = vw[idx];
    if (!R_ISFINITE(u) || !R_ISFINITE(v)) return;
    real film = R_HYPOT(u, v);
    real nlen = R_FMAX(R_HYPOT(nx, nz), K(1.0e-4));
    real into = R_FMAX((u * nx + v * nz) / nlen, K(0.0));
    real strain = R_FABS(uw[idx + 1] - uw[idx - 1]) + R_FABS(vw[idx + res] - vw[idx - res]);
    real shear_v = R_FMAX(film, strain);
    if (shear_v < K(0.2) && into < K(0.2)) return;
    real prod = R_FMAX(breaking_foam(R_FMAX(shear_v, into), h), smooth_step(K(0.25), K(1.1), into));
    *foam_out = prod > K(0.02) ? prod : K(0.0);
    if (h > K(0.012) && h < K(0.055) && (shear_v > K(0.45) || into > K(0.4))) {
        real tau = K(1000.0) * K(0.025) * shear_v * shear_v;
        if (tau > K(3.0) || into > K(0.45)) *rate_out = R_FMIN(K(0.006) * (into + shear_v * K(0.35)), K(0.01));
    }
}

static int push_wakes(const Push *p, int np) {
    for (int k = 0; k < np; k++)
        if ((R_ISFINITE(p[k].speed) && p[k].speed >= K(0.25)) || (R_ISFINITE(p[k].slip) && R_FABS(p[k].slip) >= K(0.05)))
            return 1;
    return 0;
}

/* coarse_ij (:9490-9495) */
static int coarse_ij(const Core *c, real x, real z, int *i, int *j) {
    *i = coarse_axis(c->origin[0], x);
    *j = coarse_axis(c->origin[1], z);
    return *i >= 0 && *j >= 0;
}

/* collect_tracer_wet (:6246-6290) and tracer_visit_box (:6292-6337): the
 * wet (h >= 4 mm) and shore cells of the window when a tire wakes it, else
 * of the box around the live foam, the silt and the tires' 8 m reach. */
static int collect_tracer_wet(Core *c, const Push *p, int np) {
    memset(c->wet_m, 0, WN);
    memset(c->shore_m, 0, WN);
    int i0, i1, j0, j1;
    if (push_wakes(p, np)) {
        i0 = 0; i1 = WRES - 1; j0 = 0; j1 = WRES - 1;
    } else {
        int a = WRES, b = 0, cc = WRES, d = 0, any = 0;
        for (int k = 0; k < WN; k++)
            if (c->foam_live[k]) {
                int i = k % WRES, j = k / WRES;
                if (i < a) a = i;
                if (i > b) b = i;
                if (j < cc) cc = j;
                if (j > d) d = j;
                any = 1;
            }
        int ti0, ti1, tj0, tj1;
        if (turbid_span(c->turbid, WRES, &ti0, &ti1, &tj0, &tj1)) {
            if (ti0 < a) a = ti0;
            if (ti1 > b) b = ti1;
            if (tj0 < cc) cc = tj0;
            if (tj1 > d) d = tj1;
            any = 1;
        }
        int reach = (int)R_CEIL(K(8.0) / COARSE);
        for (int k = 0; k < np; k++) {
            int ci, cj;
            if (!coarse_ij(c, p[k].x, p[k].z, &ci, &cj)) continue;
            int lo_i = ci - reach > 0 ? ci - reach : 0, hi_i = ci + reach < WRES - 1 ? ci + reach : WRES - 1;
            int lo_j = cj - reach > 0 ? cj - reach : 0, hi_j = cj + reach < WRES - 1 ? cj + reach : WRES - 1;
            if (lo_i < a) a = lo_i;
            if (hi_i > b) b = hi_i;
            if (lo_j < cc) cc = lo_j;
            if (hi_j > d) d = hi_j;
            any = 1;
        }
        if (!any) return 0;
        i0 = a >= 2 ? a - 2 : 0;
        j0 = cc >= 2 ? cc - 2 : 0;
        i1 = b + 2 < WRES - 1 ? b + 2 : WRES - 1;
        j1 = d + 2 < WRES - 1 ? d + 2 : WRES - 1;
    }
    int nwet = 0;
    for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) {
            int idx = j * WRES + i;
            real h = c->water[idx] + c->slurry[idx];
            if (!R_ISFINITE(h) || h < K(0.004)) continue;
            c->wet_m[idx] = 1;
            nwet++;
            if (shore_wet_cell(c->water, WRES, i, j)) c->shore_m[idx] = 1;
        }
    return nwet;
}

/* advect_foam (:6352-6398): semi-Lagrangian on the wet cells; a live cell
 * that is no longer wet is written 0. */
static void advect_foam(Core *c, real dt) {
    real *smp = c->tr_s;
    for (int idx = 0; idx < WN; idx++) {
        if (!c->wet_m[idx]) continue;
        real u = c->uw[idx], v = c->vw[idx];
        real s = K(0.0);
        if (R_ISFINITE(u) && R_ISFINITE(v)) {
            real x = (real)(idx % WRES) - u * dt / COARSE;
            real z = (real)(idx / WRES) - v * dt / COARSE;
            s = sample_foam(c->foam, WRES, x, z);
        }
        smp[idx] = s;
    }
    real decay = R_EXP(-dt / FOAM_LIFE_S);
    for (int idx = 0; idx < WN; idx++) {
        if (c->foam_live[idx] || c->wet_m[idx]) c->foam[idx] = K(0.0);
        if (!c->wet_m[idx]) continue;
        real v = smp[idx] * decay;
        if (v > K(1.0e-4) && R_ISFINITE(v)) c->foam[idx] = R_FMIN(v, FOAM_EQ);
    }
}

/* add_foam (:6400-6419): approach FOAM_EQ under sustained production. */
static void add_foam(Core *c, int idx, real prod, real dt) {
    if (prod <= K(0.0) || !R_ISFINITE(prod) || !(dt > K(0.0))) return;
    real cur = R_ISFINITE(c->foam[idx]) ? clampr(c->foam[idx], K(0.0), FOAM_EQ) : K(0.0);
    real room = FOAM_EQ - cur;
    if (room <= K(1.0e-5)) return;
    real grown = cur + room * prod * R_FMIN(dt / FOAM_TAU_S, K(1.0));
    c->foam[idx] = R_ISFINITE(grown) ? R_FMIN(grown, FOAM_EQ) : K(0.0);
}

/* flux_edge (:9796-9824) */
static void flux_edge(const real *h, real *rem, real *add, int a, int b, real vel, real dt) {
    if (!R_ISFINITE(vel) || R_FABS(vel) < K(1.0e-5)) return;
    int donor = vel > K(0.0) ? a : b, recv = vel > K(0.0) ? b : a;
    real column = R_ISFINITE(h[donor]) ? h[donor] : K(0.0);
    real room = column - rem[donor];
    if (room <= K(1.0e-8)) return;
    real frac = R_FMIN(R_FABS(vel) * dt / COARSE, K(0.45));
    real dh = R_FMIN(column * frac, room);
    if (dh <= K(1.0e-8) || !R_ISFINITE(dh)) return;
    rem[donor] += dh;
    add[recv] += dh;
}

/* advect_turbid (:6421-6471) with flux_thickness (:9716-9750). */
static void advect_turbid(Core *c, real dt) {
    int i0, i1, j0, j1;
    if (!turbid_span(c->turbid, WRES, &i0, &i1, &j0, &j1)) return;
    real *rem = c->tr_s, *add = c->tr_s + WN;
    for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) rem[j * WRES + i] = add[j * WRES + i] = K(0.0);
    for (int j = j0; j <= j1; j++)
        for (int i = i0; i < (i1 < WRES - 1 ? i1 : WRES - 1); i++) {
            int a = j * WRES + i;
            flux_edge(c->turbid, rem, add, a, a + 1, K(0.5) * (c->uw[a] + c->uw[a + 1]), dt);
        }
    for (int j = j0; j < (j1 < WRES - 1 ? j1 : WRES - 1); j++)
        for (int i = i0; i <= i1; i++) {
            int a = j * WRES + i;
            flux_edge(c->turbid, rem, add, a, a + WRES, K(0.5) * (c->vw[a] + c->vw[a + WRES]), dt);
        }
    for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) {
            int idx = j * WRES + i;
            if (rem[idx] == K(0.0) && add[idx] == K(0.0)) {
                if (!R_ISFINITE(c->turbid[idx])) c->turbid[idx] = K(0.0);
                continue;
            }
            real next = c->turbid[idx] - rem[idx] + add[idx];
            c->turbid[idx] = R_ISFINITE(next) ? R_FMAX(next, K(0.0)) : K(0.0);
        }
}

/* settle_turbid (:6473-6509) */
static void settle_turbid(Core *c, real dt) {
    real area = COARSE * COARSE;
    real keep = R_EXP(-dt / SILT_SETTLE_S);
    real settled = K(0.0);
    int i0, i1, j0, j1;
    if (!turbid_span(c->turbid, WRES, &i0, &i1, &j0, &j1)) return;
    for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) {
            int idx = j * WRES + i;
            real h = c->water[idx] + c->slurry[idx];
            real column = R_ISFINITE(c->turbid[idx]) ? R_FMAX(c->turbid[idx], K(0.0)) : K(0.0);
            if (column <= K(0.0)) {
                c->turbid[idx] = K(0.0);
                continue;
            }
            if (!R_ISFINITE(h) || h < K(1.0e-5)) {
                settled += column * area;
                c->turbid[idx] = K(0.0);
                continue;
            }
            real stay = column * keep;
            settled += (column - stay) * area;
            c->turbid[idx] = stay;
        }
    if (R_ISFINITE(settled)) c->turbid_settled_m3 += settled;
}

/* add_silt (:6704-6715) and lift_silt (:6691-6702) */
static void add_silt(Core *c, int idx, real dh, real area) {
    if (dh <= K(0.0) || !R_ISFINITE(dh)) return;
    real room = R_FMAX(SILT_COLUMN_CAP_M - R_FMAX(c->turbid[idx], K(0.0)), K(0.0));
    real add = R_FMIN(dh, room);
    if (add <= K(0.0)) return;
    c->turbid[idx] += add;
    c->stirred_m3 += add * area;
}

static void lift_silt(Core *c, int idx, real shear_v, real dt, real area, real rate_cap, real tau_c) {
    if (!R_ISFINITE(shear_v) || shear_v <= K(0.0)) return;
    real tau = K(1000.0) * K(0.035) * shear_v * shear_v;
    if (tau <= tau_c) return;
    real rate = R_FMIN(K(0.008) * (tau / tau_c - K(1.0)), rate_cap);
    real dh = R_FMIN(rate * dt, K(0.0025));
    add_silt(c, idx, dh, area);
}

/* spawn_tire_tracers (:6602-6669): foam on the ring where the free surface
 * cuts the tire, silt lifted by the slipping contact patch. */
static void spawn_tire_tracers(Core *c, const Push *p, real dt, real area) {
    real speed = R_ISFINITE(p->speed) ? R_FMAX(p->speed, K(0.0)) : K(0.0);
    real slip = R_ISFINITE(p->slip) ? R_FMIN(R_FABS(p->slip), K(1.5)) : K(0.0);
    if (speed < K(0.25) && slip < K(0.05)) return;
    real radius = R_FMAX(p->radius, K(0.05));
    real len = R_FMAX(R_SQRT(p->fx * p->fx + p->fz * p->fz), K(1.0e-4));
    real fx = p->fx / len, fz = p->fz / len;
    real rx = fz, rz = -fx;
    real reach = radius * K(1.15) + COARSE;
    real reach2 = reach * reach;
    real slip_v = speed * slip;
    real cell = COARSE;
    for (int idx = 0; idx < WN; idx++) {
        if (!c->wet_m[idx]) continue;
        real h = c->water[idx];
        if (!R_ISFINITE(h) || h < K(0.004)) continue;
        int i = idx % WRES, j = idx / WRES;
        real x = c->origin[0] + ((real)i + K(0.5)) * cell;
        real z = c->origin[1] + ((real)j + K(0.5)) * cell;
        real dx = x - p->x, dz = z - p->z;
        if (dx * dx + dz * dz > reach2) continue;
        real along = dx * fx + dz * fz;
        real lat = dx * rx + dz * rz;
        real dist = R_HYPOT(dx, dz);
        real film = R_HYPOT(c->uw[idx], c->vw[idx]);
        film = R_ISFINITE(film) ? film : K(0.0);
        real immerse = R_FMIN(h, radius * K(0.9));
        real r_line = R_SQRT(R_FMAX(K(2.0) * radius * immerse - immerse * immerse, K(0.0)));
        real band = K(1.0) - clampr(R_FABS(dist - r_line) / (cell * K(1.35)), K(0.0), K(1.0));
        if (band > K(0.0)) {
            real motion = R_FMAX(R_FMAX(slip_v, speed * K(0.45)), film);
            real prod = breaking_foam(motion, h) * band;
            add_foam(c, idx, prod, dt);
        }
        int on_patch = R_FABS(lat) <= p->width * K(0.55) + cell * K(0.35) && along >= -radius * K(0.55) &&
                       along <= radius * K(0.35);
        if (on_patch && h > K(0.02)) {
            real shear_v = R_FMAX(slip_v, film * K(0.35));
            lift_silt(c, idx, shear_v, dt, area, K(0.015), K(6.0));
        }
    }
}

/* spawn_foam_and_silt (:6511-6524): tires, shores, then Froude foam
 * (spawn_shore_tracers :6671-6689, spawn_breaking_speed :6542-6561). */
static void spawn_foam_and_silt(Core *c, real dt, const Push *p, int np) {
    real area = COARSE * COARSE;
    for (int k = 0; k < np; k++) spawn_tire_tracers(c, &p[k], dt, area);
    for (int idx = 0; idx < WN; idx++) {
        if (!c->shore_m[idx]) continue;
        real prod, rate;
        shore_cell(c->water, c->uw, c->vw, WRES, idx, &prod, &rate);
        add_foam(c, idx, prod, dt);
        if (rate > K(0.0)) add_silt(c, idx, R_FMIN(rate * dt, K(0.0012)), area);
    }
    for (int idx = 0; idx < WN; idx++) {
        if (!c->wet_m[idx]) continue;
        real h = c->water[idx] + c->slurry[idx];
        if (!R_ISFINITE(h) || h < K(0.004)) continue;
        real u = c->uw[idx], v = c->vw[idx];
        if (!R_ISFINITE(u) || !R_ISFINITE(v) || R_HYPOT(u, v) <= K(0.05)) continue;
        add_foam(c, idx, breaking_foam(R_HYPOT(u, v), h), dt);
    }
}

/* spawn_breaking_crest (:6563-6600): foam where the seated sheet stands
 * above the mean of its neighbours. */
static void spawn_breaking_crest(Core *c, real dt) {
    static const int nb[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    for (int idx = 0; idx < WN; idx++) {
        if (!c->wet_m[idx]) continue;
        int i = idx % WRES, j = idx / WRES;
        real h = c->water[idx] + c->slurry[idx];
        if (!R_ISFINITE(h) || h < K(0.004)) continue;
        real sum = K(0.0), count = K(0.0);
        for (int q = 0; q < 4; q++) {
            int ni = i + nb[q][0], nj = j + nb[q][1];
            if (ni < 0 || nj < 0 || ni >= WRES || nj >= WRES) continue;
            real nh = c->water[nj * WRES + ni] + c->slurry[nj * WRES + ni];
            if (R_ISFINITE(nh)) {
                sum += R_FMAX(nh, K(0.0));
                count += K(1.0);
            }
        }
        if (count <= K(0.0)) continue;
        real excess = h - sum / count;
        if (excess <= K(0.0)) continue;
        add_foam(c, idx, smooth_step(K(0.004), K(0.018), excess), dt);
    }
}

/* clear_dry_tracers (:6717-6749) */
static void clear_dry_tracers(Core *c) {
    real area = COARSE * COARSE;
    for (int idx = 0; idx < WN; idx++) {
        if (!c->wet_m[idx]) continue;
        real h = c->water[idx] + c->slurry[idx];
        if (!R_ISFINITE(h) || h < K(0.004) || !R_ISFINITE(c->foam[idx])) c->foam[idx] = K(0.0);
    }
    int i0, i1, j0, j1;
    if (!turbid_span(c->turbid, WRES, &i0, &i1, &j0, &j1)) return;
    for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) {
            int idx = j * WRES + i;
            real h = c->water[idx] + c->slurry[idx];
            if (!R_ISFINITE(h) || h < K(1.0e-5)) {
                if (c->turbid[idx] > K(0.0)) {
                    c->turbid_settled_m3 += c->turbid[idx] * area;
                    c->turbid[idx] = K(0.0);
                }
            } else if (!R_ISFINITE(c->turbid[idx])) {
                c->turbid[idx] = K(0.0);
            }
        }
}

/* sync_foam_live (:6339-6350) */
static void sync_foam_live(Core *c) {
    for (int idx = 0; idx < WN; idx++)
        c->foam_live[idx] = c->wet_m[idx] && R_ISFINITE(c->foam[idx]) && c->foam[idx] > K(1.0e-4);
}

/* step_tracers (:6210-6239) */
static void step_tracers(Core *c, real dt, const Push *p, int np) {
    if (dt <= K(0.0) || !R_ISFINITE(dt)) return;
    int wake = push_wakes(p, np);
    int live_silt = 0, live_foam = 0;
    for (int k = 0; k < WN; k++) {
        if (R_ISFINITE(c->turbid[k]) && c->turbid[k] > K(1.0e-8)) live_silt = 1;
        if (c->foam_live[k]) live_foam = 1;
    }
    if (!wake && !live_silt && !live_foam) return;
    int nwet = collect_tracer_wet(c, p, np);
    if (nwet == 0 && !live_silt) return;
    if (live_foam) advect_foam(c, dt);
    advect_turbid(c, dt);
    settle_turbid(c, dt);
    spawn_foam_and_silt(c, dt, p, np);
    clear_dry_tracers(c);
    sync_foam_live(c);
}

/* note_surface_foam (:6526-6538): crest foam after the seated sheet, from
 * the last step's pushes. */
void F(oracle_core_note_surface_foam)(Core *c, real dt) {
    if (dt <= K(0.0) || !R_ISFINITE(dt)) return;
    int nwet = collect_tracer_wet(c, c->last_pushes, c->nlast);
    if (nwet == 0) return;
    spawn_breaking_crest(c, dt);
    clear_dry_tracers(c);
    sync_foam_live(c);
}

/* Test entry point: step_tracers alone with the pushes given since the
 * last step (which it consumes, as a step does). */
void F(oracle_core_tracer_step)(Core *c, real dt) {
    step_tracers(c, dt, c->pushes, c->npush);
    memcpy(c->last_pushes, c->pushes, sizeof(Push) * (size_t)c->npush);
    c->nlast = c->npush;
    c->npush = 0;
}

/* Test entry points for the tracer laws on their own. */
real F(oracle_breaking_foam)(real speed, real thickness) { return breaking_foam(speed, thickness); }
real F(oracle_sample_foam)(const real *src, int res, real x, real z) { return sample_foam(src, res, x, z); }
void F(oracle_shore_cell)(const real *water, const real *uw, const real *vw, int res, int idx, real *out2) {
    shore_cell(water, uw, vw, res, idx, &out2[0], &out2[1]);
}
/* Lay foam and turbid planes (either may be NULL); the live foam becomes
 * the cells holding more than 1e-4 of foam. */
void F(oracle_core_set_tracers)(Core *c, const real *foam, const real *turbid) {
    if (foam) memcpy(c->foam, foam, WN * sizeof(real));
    if (turbid) memcpy(c->turbid, turbid, WN * sizeof(real));
    for (int k = 0; k < WN; k++) c->foam_live[k] = R_ISFINITE(c->foam[k]) && c->foam[k] > K(1.0e-4);
}

/* foam, turbid (200 x 200 each) and the silt accounts (stirred, settled). */
void F(oracle_core_tracers)(Core *c, real *foam, real *turbid, real *acc2) {
    if (foam) memcpy(foam, c->foam, WN * sizeof(real));
    if (turbid) memcpy(turbid, c->turbid, WN * sizeof(real));
    if (acc2) {
        acc2[0] = c->stirred_m3;
        acc2[1] = c->turbid_settled_m3;
    }
}

/* ---- suspended solids (src/fluid_core.rs:8750-9003) -------------------------
 * entrain_wake lifts solids into the water column behind each moving tire;
 * advect_suspended lets them fall out at SUSP_FALL_M_S and carries them with
 * the film.  surface_driven() (:6241-6243) is false on the sheet path (no
 * waves, no river reach), so the plume rides the along-track velocity v only
 * and there is no eddy mixing. */
#define SUSP_FALL K(0.004)

static void entrain_wake(Core *c, const Push *pushes, int np, real dt) {
    const int res = WRES;
    const real cell = COARSE, area = cell * cell;
    for (int k = 0; k < np; k++) {
        const Push *p = &pushes[k];
        if (p->speed < K(0.3)) continue;
        real fl = R_FMAX(R_SQRT(p->fx * p->fx + p->fz * p->fz), K(1.0e-4));
        real fx = p->fx / fl, fz = p->fz / fl;
        real rx = fz, rz = -fx;
        real reach = p->radius * K(3.4) + cell;
        int i0 = (int)R_FLOOR(((p->x - reach) - c->origin[0]) / cell);
        int i1 = (int)R_FLOOR(((p->x + reach) - c->origin[0]) / cell);
        int j0 = (int)R_FLOOR(((p->z - reach) - c->origin[1]) / cell);
        int j1 = (int)R_FLOOR(((p->z + reach) - c->origin[1]) / cell);
        real aft = p->radius * K(3.2);
        real half = p->width * K(0.55) + K(0.08);
        if (i0 < 0) i0 = 0;
        if (j0 < 0) j0 = 0;
        if (i1 > res - 1) i1 = res - 1;
        if (j1 > res - 1) j1 = res - 1;
        for (int j = j0; j <= j1; j++)
            for (int i = i0; i <= i1; i++) {
                real x = c->origin[0] + ((real)i + K(0.5)) * cell;
                real z = c->origin[1] + ((real)j + K(0.5)) * cell;
                real dx = x - p->x, dz = z - p->z;
                real along = dx * fx + dz * fz;
                if (along > K(-0.05) || along < -aft) continue;
                real lat = R_FABS(dx * rx + dz * rz);
                if (lat > half) continue;
                int idx = j * WRES + i;
                real water = c->water[idx];
                if (!R_ISFINITE(water) || water < K(0.004)) continue;
                real along_t = clampr((-along) / aft, K(0.0), K(1.0));
                real lat_t = clampr(lat / half, K(0.0), K(1.0));
                real core = (K(1.0) - along_t) * (K(1.0) - lat_t);
                real room = R_FMAX(water - c->susp[idx], K(0.0));
                real add = R_FMIN(K(4.0) * dt * core * water, room);
                if (add <= K(1.0e-8)) continue;
                c->susp[idx] += add;
                c->entrained_m3 += add * area;
            }
    }
}

/* film_uv (:3490-3499) of the water film alone (slurry and its velocity 0). */
static void film_uv_water(real water, real uw, real vw, real *u, real *v) {
    real h = water + K(0.0);
    *u = K(0.0);
    *v = K(0.0);
    if (!R_ISFINITE(h) || h < K(1.0e-5)) return;
    real a = (uw * water + K(0.0) * K(0.0)) / h, b = (vw * water + K(0.0) * K(0.0)) / h;
    *u = R_ISFINITE(a) ? a : K(0.0);
    *v = R_ISFINITE(b) ? b : K(0.0);
}

static void advect_suspended(Core *c, real dt) {
    const int res = WRES;
    const real cell = COARSE, area = cell * cell;
    int i0 = res, i1 = 0, j0 = res, j1 = 0;
    double landed_m3 = 0.0;
    for (int j = 0; j < res; j++)
        for (int i = 0; i < res; i++) {
            int idx = j * res + i;
            real s = c->susp[idx];
            if (!R_ISFINITE(s) || s <= K(1.0e-12)) {
                c->susp[idx] = K(0.0);
                continue;
            }
            real w = c->water[idx];
            if (!R_ISFINITE(w) || w < K(1.0e-5)) {
                c->settled_susp_m3 += s * area;
                c->susp[idx] = K(0.0);
                continue;
            }
            if (s > w) {
                c->settled_susp_m3 += s * area;
                c->susp[idx] = K(0.0);
                continue;
            }
            real keep = R_EXP(-dt * SUSP_FALL / w);
            real landed = s * (K(1.0) - keep);
            if (landed > K(0.0)) {
                landed_m3 += (double)landed * (double)area;
                s -= landed;
                c->susp[idx] = s;
            }
            if (s > K(1.0e-8)) {
                if (i < i0) i0 = i;
                if (i > i1) i1 = i;
                if (j < j0) j0 = j;
                if (j > j1) j1 = j;
            }
        }
    c->settled_susp_m3 += (real)landed_m3;
    if (i0 > i1) return;
    i0 = i0 > 0 ? i0 - 1 : 0;
    j0 = j0 > 0 ? j0 - 1 : 0;
    i1 = i1 + 1 < res - 1 ? i1 + 1 : res - 1;
    j1 = j1 + 1 < res - 1 ? j1 + 1 : res - 1;
    real *next = zalloc(WN), *taken = zalloc(WN);
    memcpy(next, c->susp, WN * sizeof(real));
    for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) {
            int idx = j * res + i;
            real depth = c->susp[idx], water = c->water[idx];
            if (depth <= K(1.0e-12) || !R_ISFINITE(depth) || !R_ISFINITE(water) || water < K(1.0e-5)) continue;
            real u, v;
            film_uv_water(water, c->uw[idx], c->vw[idx], &u, &v);
            u = K(0.0); /* not surface driven */
            if (!R_ISFINITE(v)) v = K(0.0);
            real conc = depth / water;
            real fu = R_FABS(u) * dt / cell;
            real fv = R_FABS(v) * dt / cell;
            real sum = fu + fv;
            if (sum > K(0.45)) {
                fu *= K(0.45) / sum;
                fv *= K(0.45) / sum;
            }
            int up_i = u > K(0.0) ? (i + 1 < res ? idx + 1 : -1) : (i > 0 ? idx - 1 : -1);
            int up_j = v > K(0.0) ? (j + 1 < res ? idx + res : -1) : (j > 0 ? idx - res : -1);
            const real fr[2] = {fu, fv};
            const int rc[2] = {up_i, up_j};
            for (int q = 0; q < 2; q++) {
                real give = depth * fr[q];
                if (give <= K(1.0e-12)) continue;
                int n = rc[q];
                if (n < 0) {
                    c->settled_susp_m3 += give * area;
                    next[idx] -= give;
                    continue;
                }
                real wr = c->water[n];
                if (!R_ISFINITE(wr) || wr < K(1.0e-5)) continue;
                real have = c->susp[n] + taken[n];
                real room = R_FMAX(conc * wr - have, K(0.0));
                real accept = R_FMIN(R_FMIN(give, room), R_FMAX(wr - have, K(0.0)));
                if (accept <= K(1.0e-12)) continue;
                next[idx] -= accept;
                taken[n] += accept;
            }
        }
    for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) next[j * res + i] += taken[j * res + i];
    for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++) {
            int idx = j * res + i;
            real s = next[idx];
            if (!R_ISFINITE(s) || s <= K(0.0)) {
                next[idx] = K(0.0);
                continue;
            }
            real w = c->water[idx];
            if (!R_ISFINITE(w) || w < K(1.0e-5)) {
                c->settled_susp_m3 += s * area;
                s = K(0.0);
            } else if (s > w) {
                c->settled_susp_m3 += (s - w) * area;
                s = w;
            }
            next[idx] = s;
        }
    memcpy(c->susp, next, WN * sizeof(real));
    free(next);
    free(taken);
}

/* Test entry points: entrain_wake of the pushes given, and step_suspended. */
void F(oracle_core_entrain)(Core *c, real dt) { entrain_wake(c, c->pushes, c->npush, dt); c->npush = 0; }
void F(oracle_core_suspended_step)(Core *c, real dt) {
    if (dt > K(0.0) && R_ISFINITE(dt)) advect_suspended(c, dt);
}
/* The suspended plane (200 x 200) in or out, and the accounts: [0] column
 * (column_sum of the plane), [1] settled, [2] parked off the window,
 * [3] entrained. */
void F(oracle_core_set_susp)(Core *c, const real *s) { memcpy(c->susp, s, WN * sizeof(real)); }
void F(oracle_core_susp)(Core *c, real *s, real *acc4) {
    if (s) memcpy(s, c->susp, WN * sizeof(real));
    if (acc4) {
        acc4[0] = F(oracle_column_sum)(c->susp, WN, COARSE);
        acc4[1] = c->settled_susp_m3;
        real fz = K(0.0);
        for (size_t k = 0; k < c->pcap; k++)
            if (c->parked[k].state == 1) fz += c->parked[k].susp * COARSE * COARSE;
        acc4[2] = fz;
        acc4[3] = c->entrained_m3;
    }
}

/* step_pair (src/fluid_core.rs:7942-8150), no lifts or reach.  With walls
 * the mask of this grid is laid once from the planes before either layer
 * steps (:7954-7966). */
static void step_pair(Core *c, real dt, int fine) {
    if (fine && !c->fine_on) return;
    if (c->nsolids > 0) {
        uint8_t *mask = fine ? c->fmask : c->mask;
        solid_mask(c, fine, mask);
        step_layer_walled(c, fine, 0, mask, dt);
        int res = fine ? FRES : WRES;
        fill_nu_s(c, fine ? c->fs : c->slurry, fine ? c->fmoist : c->moisture, fine ? c->fgrain : c->grain, res);
        step_layer_walled(c, fine, 1, mask, dt);
        return;
    }
    int res = fine ? FRES : WRES;
    real cell = fine ? FINE : COARSE;
    real cap = fine ? c->fine_sub_cap : K(8.0);
    int pace = fine ? c->fine_pace : 0;
    int confine = fine ? 0 : c->coarse_confine;
    real *bed = fine ? c->fbed : c->bed;
    real *bed_s = fine ? c->fbed_s : c->bed;
    if (!fine) soften_border(c->uw, c->vw, res);
    F(oracle_step_sheet_ex)(fine ? c->fw : c->water, fine ? c->fuw : c->uw, fine ? c->fvw : c->vw, bed, c->nu_w,
                            c->gate, res, cell, dt, 0, K(0.0), K(0.0), K(0.08), K(1.0e-6), 0, cap, pace, 0,
                            confine, NULL, NULL);
    fill_nu_s(c, fine ? c->fs : c->slurry, fine ? c->fmoist : c->moisture, fine ? c->fgrain : c->grain, res);
    if (!fine) soften_border(c->us, c->vs, res);
    F(oracle_step_sheet_ex)(fine ? c->fs : c->slurry, fine ? c->fus : c->us, fine ? c->fvs : c->vs, bed_s, c->nu_s,
                            c->gate, res, cell, dt, 0, K(0.0), K(0.0), K(0.05), K(1.0e-6), 0, cap, pace, 0,
                            confine, NULL, NULL);
}

/* ---- public oracle API ----------------------------------------------------- */

/* Test hooks: the two halves of the patch hand-off on their own
 * (near.rs:384-410 calls refine and restrict_fine directly). */
void F(oracle_core_refine)(Core *c, real dt) { refine(c, c->pushes, c->npush, dt); }
void F(oracle_core_restrict)(Core *c) { restrict_fine(c); }
real F(oracle_floor_scale)(real base, const real *r) { return floor_scale(base, r); }
void F(oracle_recentre)(real *r) { recentre(r); }
int F(oracle_core_fine_on)(Core *c) { return c->fine_on; }
void F(oracle_core_fine_get)(Core *c, real *fw, real *origin2) {
    memcpy(fw, c->fw, FN * sizeof(real));
    origin2[0] = c->fine_origin[0];
    origin2[1] = c->fine_origin[1];
}

/* fine_water_debug (near.rs:341-346): lay the patch under the given pushes
 * with dt = 0, read it, then restrict it again.  fuw and fvw may be NULL. */
void F(oracle_core_fine_water_debug)(Core *c, const real *p8, int np, real *fw, real *fuw, real *fvw,
                                     real *origin2) {
    Push q[MAX_PUSH];
    if (np > MAX_PUSH) np = MAX_PUSH;
    for (int k = 0; k < np; k++) {
        const real *p = p8 + 8 * k;
        Push one = {p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]};
        q[k] = one;
    }
    refine(c, q, np, K(0.0));
    memcpy(fw, c->fw, FN * sizeof(real));
    if (fuw) memcpy(fuw, c->fuw, FN * sizeof(real));
    if (fvw) memcpy(fvw, c->fvw, FN * sizeof(real));
    origin2[0] = c->fine_origin[0];
    origin2[1] = c->fine_origin[1];
    restrict_fine(c);
}

/*
 * Test entry points: push_on, peel_spray and soften_border on their own.
 * p8: x, z, forward x, forward z, radius, width, speed, slip.
 * oracle_core_peel uses the pushes given with oracle_core_push_tire and
 * returns the volume peeled; events receives the event count.
 */
void F(oracle_core_push_on)(Core *c, const real *p8, int fine) {
    Push p = {p8[0], p8[1], p8[2], p8[3], p8[4], p8[5], p8[6], p8[7]};
    push_on(c, &p, fine);
}
real F(oracle_core_peel)(Core *c, int fine, int *events) {
    *events = 0;
    return peel_spray(c, fine, 0, events);
}
void F(oracle_soften_border)(real *u, real *v, int res) { soften_border(u, v, res); }

void F(oracle_core_push_tire)(Core *c, const real *p) {
    if (c->npush >= MAX_PUSH) return;
    Push q = {p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]};
    c->pushes[c->npush++] = q;
}

/* FluidCore::step on the sheet path.  Returns the spray events; *spray the
 * volume peeled this step. */
int F(oracle_core_step)(Core *c, real dt, real *spray) {
    Push pushes[MAX_PUSH];
    int np = c->npush;
    memcpy(pushes, c->pushes, sizeof(Push) * (size_t)np);
    int events = 0;
    real vol = K(0.0);
    /* lift from the column the tire rolls into, before the sheet shoves it (:6050-6053) */
    if (dt > K(0.0) && R_ISFINITE(dt) && np > 0) entrain_wake(c, pushes, np, dt);
    refine(c, pushes, np, dt);
    for (int k = 0; k < np; k++)
        if (cell_is_refined(c, pushes[k].x, pushes[k].z)) push_on(c, &pushes[k], 1);
    step_pair(c, dt, 1);
    if (c->fine_on) vol += peel_spray(c, 1, 0, &events);
    restrict_fine(c);
    for (int k = 0; k < np; k++)
        if (!fine_held_at(c, pushes[k].x, pushes[k].z)) push_on(c, &pushes[k], 0);
    step_pair(c, dt, 0);
    vol += peel_spray(c, 0, 1, &events);
    if (c->nsolids > 0) scour_contacts(c, dt);
    step_tracers(c, dt, pushes, np);
    /* step_suspended (:8750-8756) */
    if (dt > K(0.0) && R_ISFINITE(dt)) advect_suspended(c, dt);
    memcpy(c->last_pushes, pushes, sizeof(Push) * (size_t)np);
    c->nlast = np;
    c->npush = 0;
    if (spray) *spray = vol;
    return events;
}


/* set_hydro_solids (:5257-5260): s10 holds n solids of x, z, yaw,
 * half_length, half_width, bottom_y, top_y, contact, radius, width. */
void F(oracle_core_set_hydro_solids)(Core *c, const real *s10, int n) {
    if (n > MAX_SOLIDS) n = MAX_SOLIDS;
    if (n < 0) n = 0;
    for (int k = 0; k < n; k++) {
        const real *q = s10 + 10 * k;
        HSolid h = {q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7] != K(0.0), q[8], q[9]};
        c->solids[k] = h;
    }
    c->nsolids = n;
}

/* The wall mask of the window (fine 0) or the patch (fine 1). */
void F(oracle_core_solid_mask)(Core *c, int fine, uint8_t *out) { solid_mask(c, fine, out); }

/* hydro_load (:7164-7249).  out10: force_x, force_z, buoyancy_n,
 * hull_buoyancy_n, wheel_buoyancy_n, displaced_m3, eroded_m3, hole_m,
 * wheel_blocks, belly_blocks. */
void F(oracle_core_hydro_load)(Core *c, real *out10) {
    uint8_t *mask = c->mask;
    if (c->nsolids == 0) memset(mask, 0, WN);
    else solid_mask(c, 0, mask);
    real fx = K(0.0), fz = K(0.0);
    unsigned wheel_blocks = 0, belly_blocks = 0;
    for (size_t k = 0; k < WN; k++) {
        if (mask[k] == 1) wheel_blocks++;
        else if (mask[k] == 2) belly_blocks++;
    }
    static const int nb[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    for (int j = 0; j < WRES; j++)
        for (int i = 0; i < WRES; i++) {
            size_t idx = (size_t)j * WRES + i;
            if (mask[idx] != 0) continue;
            real h = c->water[idx];
            if (h < K(1.0e-4) || !R_ISFINITE(h)) continue;
            for (int q = 0; q < 4; q++) {
                int ni = i + nb[q][0], nj = j + nb[q][1];
                if (ni < 0 || nj < 0 || ni >= WRES || nj >= WRES) continue;
                if (mask[(size_t)nj * WRES + ni] == 0) continue;
                real nx = (real)nb[q][0], nz = (real)nb[q][1];
                real hydro = K(1000.0) * GRAV * h * h * K(0.5) * COARSE;
                fx += nx * hydro;
                fz += nz * hydro;
                real vn = c->uw[idx] * nx + c->vw[idx] * nz;
                if (vn > K(0.0)) {
                    real dyn_f = K(0.5) * K(1000.0) * vn * vn * h * COARSE;
                    fx += nx * dyn_f;
                    fz += nz * dyn_f;
                }
            }
        }
    real wheel_b = K(0.0), hull_b = K(0.0), displaced = K(0.0);
    for (int k = 0; k < c->nsolids; k++) {
        const HSolid *s = &c->solids[k];
        real bed, water;
        if (!ring_surface(c, s, &bed, &water)) continue;
        if (s->radius > K(0.0)) {
            real vol = cylinder_segment_volume(water, s->radius, s->width);
            displaced += vol;
            wheel_b += K(1000.0) * GRAV * vol;
        } else {
            real eta = bed + water;
            real height = R_FMAX(s->top_y - s->bottom_y, K(0.0));
            real sub = clampr(eta - s->bottom_y, K(0.0), height);
            real vol = (s->half_length * K(2.0)) * (s->half_width * K(2.0)) * sub * K(0.36);
            displaced += vol;
            hull_b += K(1000.0) * GRAV * vol;
        }
    }
    out10[0] = fx; out10[1] = fz; out10[2] = wheel_b + hull_b; out10[3] = hull_b; out10[4] = wheel_b;
    out10[5] = displaced; out10[6] = c->eroded_m3; out10[7] = c->scour_hole_m;
    out10[8] = (real)wheel_blocks; out10[9] = (real)belly_blocks;
}

/* Test entry point: scour_contacts on its own. */
void F(oracle_core_scour)(Core *c, real dt) { scour_contacts(c, dt); }

/* The scour's bed drop plane (200 x 200). */
void F(oracle_core_bed_drop)(Core *c, real *out) { memcpy(out, c->bed_drop, WN * sizeof(real)); }

/* map_water_bed with arrays sampled at cell centres (src/fluid_core.rs:5329-5355). */
void F(oracle_core_map_water_bed)(Core *c, const real *bed, const real *water) {
    restrict_fine(c);
    for (int k = 0; k < WN; k++) {
        c->bed[k] = isfinite(bed[k]) ? bed[k] : K(0.0);
        real w = isfinite(water[k]) ? R_FMAX(water[k], K(0.0)) : K(0.0);
        c->water[k] = w;
        if (w > K(0.0)) {
            c->moisture[k] = R_FMAX(c->moisture[k], K(0.8));
            if (c->grain[k] == 4) c->grain[k] = 1;
        }
    }
}

/* paint (src/fluid_core.rs:5820-5855): pour_water / pour_slurry. */
void F(oracle_core_paint)(Core *c, real x0, real x1, real z0, real z1, real depth, int water, int grain,
                          real moisture) {
    if (!(depth > K(0.0))) depth = K(0.0);
    for (int j = 0; j < WRES; j++)
        for (int i = 0; i < WRES; i++) {
            real x, z;
            coarse_xz(c, i, j, &x, &z);
            if (x < x0 || x >= x1 || z < z0 || z >= z1) continue;
            int idx = j * WRES + i;
            if (water) {
                c->water[idx] = depth;
            } else {
                c->slurry[idx] = depth;
                c->grain[idx] = (uint8_t)grain;
                c->moisture[idx] = moisture < K(0.0) ? K(0.0) : (moisture > K(1.0) ? K(1.0) : moisture);
            }
        }
}

/* lay_slurry_mask with a depth array (src/fluid_core.rs:5857-5880). */
void F(oracle_core_lay_slurry)(Core *c, const real *depth, int grain, real moisture) {
    restrict_fine(c);
    real m = moisture < K(0.0) ? K(0.0) : (moisture > K(1.0) ? K(1.0) : moisture);
    for (int k = 0; k < WN; k++) {
        real d = depth[k];
        if (!(isfinite(d) && d > K(0.0))) continue;
        c->slurry[k] += d;
        c->grain[k] = (uint8_t)grain;
        c->moisture[k] = R_FMAX(c->moisture[k], m);
    }
}

/* read_substrate_grid with arrays (src/fluid_core.rs:5928-5962), with the
 * ambient moisture histogram (1 % buckets, half away from zero). */
void F(oracle_core_read_substrate)(Core *c, const real *bed, const real *moisture, const uint8_t *grain) {
    restrict_fine(c);
    unsigned hist[101];
    memset(hist, 0, sizeof hist);
    for (int k = 0; k < WN; k++) {
        real m = moisture[k];
        m = m < K(0.0) ? K(0.0) : (m > K(1.0) ? K(1.0) : m);
        c->bed[k] = bed[k];
        c->moisture[k] = m;
        float pct = (float)m * 100.0f;
        int whole = (int)pct;
        int bucket = pct - (float)whole >= 0.5f ? whole + 1 : whole;
        hist[bucket < 100 ? bucket : 100]++;
        if (c->slurry[k] > K(1.0e-5)) c->grain[k] = grain[k];
    }
    int best = 0;
    unsigned best_n = 0;
    for (int k = 0; k <= 100; k++)
        if (hist[k] > best_n) {
            best = k;
            best_n = hist[k];
        }
    c->moisture_ambient = (real)((float)best * 0.01f);
}

/* Copies the coarse planes out (water, slurry, uw, vw, us, vs, bed, moisture). */
void F(oracle_core_get)(Core *c, real *out8) {
    real *planes[8] = {c->water, c->slurry, c->uw, c->vw, c->us, c->vs, c->bed, c->moisture};
    for (int p = 0; p < 8; p++) memcpy(out8 + (size_t)p * WN, planes[p], WN * sizeof(real));
}

void F(oracle_core_set)(Core *c, const real *in8) {
    real *planes[8] = {c->water, c->slurry, c->uw, c->vw, c->us, c->vs, c->bed, c->moisture};
    for (int p = 0; p < 8; p++) memcpy(planes[p], in8 + (size_t)p * WN, WN * sizeof(real));
}

void F(oracle_core_origin)(Core *c, real *out2) {
    out2[0] = c->origin[0];
    out2[1] = c->origin[1];
}

/* ledger() for the sheet path (src/fluid_core.rs:6821-6883): out[0] water,
 * [1] slurry, [2] frozen water, [3] frozen slurry, [4] spray, [5] total. */
void F(oracle_core_ledger)(Core *c, real *out) {
    real area = COARSE * COARSE;
    real w = -K(0.0), s = -K(0.0);
    for (int k = 0; k < WN; k++) {
        w += R_FMAX(c->water[k], K(0.0));
        s += R_FMAX(c->slurry[k], K(0.0));
    }
    w *= area;
    s *= area;
    if (c->fine_on) {
        real fa = FINE * FINE, fw = -K(0.0), fs = -K(0.0);
        for (int k = 0; k < FN; k++) {
            fw += R_FMAX(c->fw[k], K(0.0));
            fs += R_FMAX(c->fs[k], K(0.0));
        }
        w += fw * fa;
        s += fs * fa;
    }
    out[0] = w;
    out[1] = s;
    double pw = 0.0, ps = 0.0;
    for (size_t k = 0; k < c->pcap; k++)
        if (c->parked[k].state == 1) {
            pw += (double)c->parked[k].water;
            ps += (double)c->parked[k].slurry;
        }
    out[2] = c->frozen_w + (real)(pw * (double)area);
    out[3] = c->frozen_s + (real)(ps * (double)area);
    out[4] = c->spray_m3;
    out[5] = w + s + out[2] + out[3] + c->spray_m3;
}

/* The render-contract film and flow maps from coarse_cell
 * (src/fluid_core.rs:6998-7034; src/render_contract.rs:468-496):
 * film2 = (water, slurry), flow4 = (u, v, speed, turbidity). */
static real finite0r(real x) { return isfinite((double)x) ? x : K(0.0); }

void F(oracle_core_publish)(Core *c, real *film2, real *flow4) {
    for (int k = 0; k < WN; k++) {
        real water = isfinite(c->water[k]) ? c->water[k] : K(0.0);
        real slurry = isfinite(c->slurry[k]) ? c->slurry[k] : K(0.0);
        real film = water + slurry;
        real u = K(0.0), v = K(0.0);
        if (!(film < K(1.0e-5))) {
            u = (c->uw[k] * water + c->us[k] * slurry) / film;
            v = (c->vw[k] * water + c->vs[k] * slurry) / film;
            if (!isfinite(u)) u = K(0.0);
            if (!isfinite(v)) v = K(0.0);
        }
        int dry = !isfinite(water) || !isfinite(slurry) || water + slurry < K(1.0e-5);
        real speed = dry ? K(0.0) : R_SQRT(u * u + v * v);
        if (!isfinite(speed)) speed = K(0.0);
        film2[2 * k] = water;
        film2[2 * k + 1] = slurry;
        flow4[4 * k] = dry ? K(0.0) : u;
        flow4[4 * k + 1] = dry ? K(0.0) : v;
        flow4[4 * k + 2] = speed;
        /* coarse_cell's turbidity (:7025-7032): wheel-lifted solids plus
         * stirred silt over the water; the contract's channel (:476-481) */
        real turb = K(0.0);
        if (!(water < K(1.0e-5))) {
            real load = R_FMAX(finite0r(c->susp[k]), K(0.0)) + R_FMAX(finite0r(c->turbid[k]), K(0.0));
            turb = finite0r(clampr(load / water, K(0.0), K(1.0)));
            turb = clampr(turb, K(0.0), K(1.0));
        }
        flow4[4 * k + 3] = turb;
    }
}

/* The render contract's moisture, wet_edge and foam_crest maps
 * (src/render_contract.rs:457-500 and :526-569) over the window: moisture =
 * max(coarse_cell moisture, soil wetness) (wetness NULL: no soil texel);
 * wet = moisture >= 0.45 or water + slurry >= 0.004 (is_wet, :261-263);
 * wet_edge = 1 where an in-window 4-neighbour's wet differs; foam_crest =
 * (foam, placed + excess) with excess = max(film - mean of the in-window
 * 4-neighbours' films, 0) and placed = max(ground_rut placed channel, 0)
 * (placed NULL: 0).  No foam tracer here: foam is 0. */

void F(oracle_core_publish_surface)(Core *c, const real *wetness, const real *placed, real *moisture1,
                                    real *wet_edge1, real *foam2) {
    uint8_t *wet = (uint8_t *)malloc(WN);
    real *film = (real *)malloc(sizeof(real) * WN);
    for (int k = 0; k < WN; k++) {
        real water = finite0r(c->water[k]), slurry = finite0r(c->slurry[k]);
        real m = finite0r(c->moisture[k]);
        if (wetness) m = R_FMAX(m, wetness[k]);
        moisture1[k] = finite0r(m);
        wet[k] = m >= K(0.45) || water + slurry >= K(0.004);
        film[k] = water + slurry;
    }
    static const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
    for (int j = 0; j < WRES; j++)
        for (int i = 0; i < WRES; i++) {
            int idx = j * WRES + i;
            real edge = K(0.0);
            for (int q = 0; q < 4; q++) {
                int ni = i + di[q], nj = j + dj[q];
                if (ni < 0 || nj < 0 || ni >= WRES || nj >= WRES) continue;
                if (wet[idx] != wet[nj * WRES + ni]) {
                    edge = K(1.0);
                    break;
                }
            }
            wet_edge1[idx] = edge;
            real h = film[idx], sum = K(0.0), count = K(0.0);
            for (int q = 0; q < 4; q++) {
                int ni = i + di[q], nj = j + dj[q];
                if (ni < 0 || nj < 0 || ni >= WRES || nj >= WRES) continue;
                sum += film[nj * WRES + ni];
                count += K(1.0);
            }
            real mean = count > K(0.0) ? sum / count : h;
            real excess = finite0r(R_FMAX(h - mean, K(0.0)));
            real pl = placed ? finite0r(R_FMAX(placed[idx], K(0.0))) : K(0.0);
            /* foam_crest (render_contract.rs:563-568): the foam tracer
             * where the film is at least 4 mm. */
            real fm = h >= K(0.004) ? clampr(finite0r(c->foam[idx]), K(0.0), K(1.0)) : K(0.0);
            foam2[2 * idx] = fm;
            foam2[2 * idx + 1] = finite0r(pl + excess);
        }
    free(film);
    free(wet);
}

/* drain_over_bed (src/fluid_core.rs:5123-5214) with step_water_drain /
 * step_slurry_drain (:5159-5214) and restore_wiped (:4622-4637). */
static void restore_wiped(real *h, real before) {
    real lost = before - F(oracle_column_sum)(h, WN, COARSE);
    if (lost <= K(1.0e-8)) return;
    int best = -1;
    for (int k = 0; k < WN; k++)
        if (isfinite((double)h[k]) && (best < 0 || !(h[k] < h[best]))) best = k;
    if (best >= 0) h[best] += lost / (COARSE * COARSE);
}

static real peak_speed(const real *h, const real *u, const real *v) {
    real peak = K(0.0);
    for (int k = 0; k < WN; k++)
        if (h[k] > K(1.0e-4)) {
            real a = u[k] < 0 ? -u[k] : u[k], b = v[k] < 0 ? -v[k] : v[k];
            if (a > peak) peak = a;
            if (b > peak) peak = b;
        }
    return peak;
}

real F(oracle_core_drain_over_bed)(Core *c, real dt, unsigned steps) {
    if (!(isfinite((double)dt) && dt > K(0.0)) || steps == 0) return F(oracle_column_sum)(c->water, WN, COARSE);
    restrict_fine(c);
    real *nu_w = zalloc(WN), *nu_s = zalloc(WN);
    uint8_t *gate = (uint8_t *)malloc(WN);
    memset(gate, 1, WN);
    for (int k = 0; k < WN; k++) nu_w[k] = F(oracle_dynamic_viscosity)(3, K(1.0));
    unsigned calm = 0;
    int moved = 0;
    for (unsigned it = 0; it < steps; it++) {
        real wb = F(oracle_column_sum)(c->water, WN, COARSE);
        real sb = F(oracle_column_sum)(c->slurry, WN, COARSE);
        soften_border(c->uw, c->vw, WRES);
        F(oracle_step_sheet_ex)(c->water, c->uw, c->vw, c->bed, nu_w, gate, WRES, COARSE, dt, 0, K(0.0), K(0.0),
                                K(0.0), K(1.0e-6), 1, K(8.0), 2, 1, 0, NULL, NULL);
        for (int k = 0; k < WN; k++) nu_s[k] = F(oracle_dynamic_viscosity)(c->grain[k], c->moisture[k]);
        soften_border(c->us, c->vs, WRES);
        F(oracle_step_sheet_ex)(c->slurry, c->us, c->vs, c->bed, nu_s, gate, WRES, COARSE, dt, 0, K(0.0), K(0.0),
                                K(0.0), K(1.0e-6), 1, K(8.0), 2, 1, 0, NULL, NULL);
        real keep = (real)expf((float)(K(-0.55) * dt));
        for (int k = 0; k < WN; k++) {
            c->uw[k] *= keep;
            c->vw[k] *= keep;
            c->us[k] *= keep;
            c->vs[k] *= keep;
        }
        restore_wiped(c->water, wb);
        restore_wiped(c->slurry, sb);
        real sw = peak_speed(c->water, c->uw, c->vw), ss = peak_speed(c->slurry, c->us, c->vs);
        real speed = sw > ss ? sw : ss;
        if (speed > K(0.05)) moved = 1;
        if (moved && speed < K(0.008)) {
            if (++calm >= 24) break;
        } else {
            calm = 0;
        }
    }
    free(nu_w);
    free(nu_s);
    free(gate);
    return F(oracle_column_sum)(c->water, WN, COARSE);
}

/* ---- recenter (src/fluid_core.rs:5409-5422, shift_coarse :9004-9226) ------- */

static size_t park_hash(int32_t kx, int32_t kz, size_t cap) {
    uint64_t h = ((uint64_t)(uint32_t)kx * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)(uint32_t)kz * 0xC2B2AE3D27D4EB4Full);
    return (size_t)(h ^ (h >> 29)) & (cap - 1);
}

static struct ParkedSlot *park_find(Core *c, int32_t kx, int32_t kz) {
    if (!c->pcap) return NULL;
    for (size_t k = park_hash(kx, kz, c->pcap), n = 0; n < c->pcap; k = (k + 1) & (c->pcap - 1), n++) {
        struct ParkedSlot *p = &c->parked[k];
        if (p->state == 0) return NULL;
        if (p->state == 1 && p->kx == kx && p->kz == kz) return p;
    }
    return NULL;
}

static void park_grow(Core *c) {
    size_t cap = c->pcap ? c->pcap * 2 : 1u << 16;
    struct ParkedSlot *old = c->parked;
    size_t ocap = c->pcap;
    c->parked = (struct ParkedSlot *)calloc(cap, sizeof(struct ParkedSlot));
    c->pcap = cap;
    c->pcount = c->pused = 0;
    for (size_t k = 0; k < ocap; k++)
        if (old[k].state == 1) {
            size_t j = park_hash(old[k].kx, old[k].kz, cap);
            while (c->parked[j].state) j = (j + 1) & (cap - 1);
            c->parked[j] = old[k];
            c->pcount++;
            c->pused++;
        }
    free(old);
}

static void park_put(Core *c, int32_t kx, int32_t kz, const struct ParkedSlot *v) {
    struct ParkedSlot *p = park_find(c, kx, kz);
    if (!p) {
        if ((c->pused + 1) * 2 > c->pcap) park_grow(c);
        size_t k = park_hash(kx, kz, c->pcap);
        while (c->parked[k].state == 1) k = (k + 1) & (c->pcap - 1);
        p = &c->parked[k];
        if (p->state == 0) c->pused++;
        c->pcount++;
    }
    *p = *v;
    p->kx = kx;
    p->kz = kz;
    p->state = 1;
}

static void park_remove(Core *c, struct ParkedSlot *p) {
    p->state = 2;
    c->pcount--;
}

static void world_key_c(real x, real z, int32_t *kx, int32_t *kz) {
    *kx = (int32_t)R_FLOOR(x / COARSE);
    *kz = (int32_t)R_FLOOR(z / COARSE);
}

static void park_cell(Core *c, const real *org, int i, int j, real *const *h) {
    int idx = j * WRES + i;
    int32_t kx, kz;
    world_key_c(org[0] + ((real)i + K(0.5)) * COARSE, org[1] + ((real)j + K(0.5)) * COARSE, &kx, &kz);
    struct ParkedSlot v = {0, 0, 1, h[0][idx], h[1][idx], h[2][idx], h[3][idx], h[4][idx], h[5][idx], h[6][idx],
                           h[7][idx], h[8][idx], h[9][idx]};
    real dm = v.moisture - c->moisture_ambient;
    if (dm < 0) dm = -dm;
    /* park_slot (:4797-4805) */
    if (v.water > K(1.0e-8) || v.slurry > K(1.0e-8) || v.susp > K(1.0e-8) || dm > K(6.0e-3) ||
        v.foam > K(1.0e-4) || v.turbid > K(1.0e-8)) {
        park_put(c, kx, kz, &v);
    } else {
        struct ParkedSlot *p = park_find(c, kx, kz);
        if (p) park_remove(c, p);
    }
}

static void restore_cell(Core *c, int i, int j, real *const *o) {
    int idx = j * WRES + i;
    int32_t kx, kz;
    world_key_c(c->origin[0] + ((real)i + K(0.5)) * COARSE, c->origin[1] + ((real)j + K(0.5)) * COARSE, &kx, &kz);
    struct ParkedSlot *p = park_find(c, kx, kz);
    if (!p) return;
    const real v[10] = {p->water, p->slurry, p->uw, p->vw, p->us, p->vs, p->moisture, p->foam, p->turbid, p->susp};
    for (int k = 0; k < 10; k++) o[k][idx] = v[k];
    park_remove(c, p);
}

typedef void (*cell_fn)(Core *, const real *, int, int, real *const *);

static void park_cb(Core *c, const real *org, int i, int j, real *const *h) { park_cell(c, org, i, j, h); }
static void restore_cb(Core *c, const real *org, int i, int j, real *const *o) {
    (void)org;
    restore_cell(c, i, j, o);
}

static void each_leaving_c(Core *c, int di, int dj, cell_fn f, const real *org, real *const *planes) {
    const int n = WRES;
    if (di > 0) {
        int w = di < n ? di : n;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < w; i++) f(c, org, i, j, planes);
    } else if (di < 0) {
        int w = -di < n ? -di : n;
        for (int j = 0; j < n; j++)
            for (int i = n - w; i < n; i++) f(c, org, i, j, planes);
    }
    if (dj != 0) {
        int h = (dj > 0 ? dj : -dj) < n ? (dj > 0 ? dj : -dj) : n;
        int jlo = dj > 0 ? 0 : n - h, jhi = dj > 0 ? h : n;
        for (int j = jlo; j < jhi; j++)
            for (int i = 0; i < n; i++) {
                int covered = di > 0 ? i < di : (di < 0 ? i >= n + di : 0);
                if (!covered) f(c, org, i, j, planes);
            }
    }
}

static int uniform_shift_c(const real *o, const real *nw, int *di, int *dj) {
    int last = WRES - 1, mid = WRES / 2;
    int32_t o0x, o0z, n0x, n0z, o1x, o1z, n1x, n1z, omx, omz, nmx, nmz;
    world_key_c(o[0] + K(0.5) * COARSE, o[1] + K(0.5) * COARSE, &o0x, &o0z);
    world_key_c(nw[0] + K(0.5) * COARSE, nw[1] + K(0.5) * COARSE, &n0x, &n0z);
    *di = n0x - o0x;
    *dj = n0z - o0z;
    world_key_c(o[0] + ((real)last + K(0.5)) * COARSE, o[1] + ((real)last + K(0.5)) * COARSE, &o1x, &o1z);
    world_key_c(nw[0] + ((real)last + K(0.5)) * COARSE, nw[1] + ((real)last + K(0.5)) * COARSE, &n1x, &n1z);
    if (o1x - o0x != last || o1z - o0z != last) return 0;
    if (n1x - n0x != last || n1z - n0z != last) return 0;
    if (n1x - o1x != *di || n1z - o1z != *dj) return 0;
    world_key_c(o[0] + ((real)mid + K(0.5)) * COARSE, o[1] + ((real)mid + K(0.5)) * COARSE, &omx, &omz);
    world_key_c(nw[0] + ((real)mid + K(0.5)) * COARSE, nw[1] + ((real)mid + K(0.5)) * COARSE, &nmx, &nmz);
    return nmx - omx == *di && nmz - omz == *dj;
}

void F(oracle_core_recenter)(Core *c, real x, real z) {
    restrict_fine(c);
    real extent = (real)WRES * COARSE;
    real nw[2] = {R_FLOOR((x - extent * K(0.5)) / COARSE) * COARSE, R_FLOOR((z - extent * K(0.5)) / COARSE) * COARSE};
    real d0 = nw[0] - c->origin[0], d1 = nw[1] - c->origin[1];
    if ((d0 < 0 ? -d0 : d0) < K(1.0e-4) && (d1 < 0 ? -d1 : d1) < K(1.0e-4)) return;
    real *h[10] = {c->water, c->slurry, c->uw, c->vw, c->us, c->vs, c->moisture, c->foam, c->turbid, c->susp};
    real *o[10];
    for (int k = 0; k < 10; k++) {
        o[k] = zalloc(WN);
        if (k == 6)
            for (int q = 0; q < WN; q++) o[k][q] = c->moisture_ambient;
    }
    int di, dj;
    real old[2] = {c->origin[0], c->origin[1]};
    if (uniform_shift_c(c->origin, nw, &di, &dj)) {
        each_leaving_c(c, di, dj, park_cb, old, h);
        for (int j = 0; j < WRES; j++)
            for (int i = 0; i < WRES; i++) {
                int si = i + di, sj = j + dj;
                if (si < 0 || sj < 0 || si >= WRES || sj >= WRES) continue;
                int src = sj * WRES + si, dst = j * WRES + i;
                if (!(h[0][src] > K(1.0e-8) || h[1][src] > K(1.0e-8) || h[6][src] > K(1.0e-4) ||
                      h[7][src] > K(1.0e-5) || h[8][src] > K(1.0e-8) || h[9][src] > K(1.0e-8)))
                    continue;
                for (int k = 0; k < 10; k++) o[k][dst] = h[k][src];
            }
        c->origin[0] = nw[0];
        c->origin[1] = nw[1];
        each_leaving_c(c, -di, -dj, restore_cb, c->origin, o);
    } else {
        for (int j = 0; j < WRES; j++)
            for (int i = 0; i < WRES; i++) park_cell(c, old, i, j, h);
        c->origin[0] = nw[0];
        c->origin[1] = nw[1];
        for (int j = 0; j < WRES; j++)
            for (int i = 0; i < WRES; i++) restore_cell(c, i, j, o);
    }
    for (int k = 0; k < 10; k++) {
        memcpy(h[k], o[k], WN * sizeof(real));
        free(o[k]);
    }
    /* reset_tracer_indices (:9228-9233) */
    memset(c->foam_live, 0, WN);
}

long F(oracle_core_parked_cells)(Core *c) { return (long)c->pcount; }

/* Parked volumes (cubic metres) for the ledger's frozen lines. */
void F(oracle_core_parked_volume)(Core *c, double *out2) {
    double w = 0.0, sl = 0.0;
    for (size_t k = 0; k < c->pcap; k++)
        if (c->parked[k].state == 1) {
            w += (double)c->parked[k].water;
            sl += (double)c->parked[k].slurry;
        }
    out2[0] = w * (double)COARSE * (double)COARSE;
    out2[1] = sl * (double)COARSE * (double)COARSE;
}

/* ------------------------------------------------------------------------
 * Ponds in tire cuts: pond_ruts (src/fluid_core.rs:5509-5611),
 * seat_rut_water (:5622-5736) and shove_rut_slurry (:5740-5792), with the
 * summed-area helpers prefix_plane / plane_window (:1502-1525).
 * ------------------------------------------------------------------------ */

/* prefix_plane (:1502-1516): inclusive 2D prefix into a (n+1)^2 plane whose
 * first row and column stay zero; each row is a running sum along the row,
 * added to the prefix row above. */
static void prefix_plane(const real *src, int n, real *dst) {
    int stride = n + 1;
    for (int j = 0; j < n; j++) {
        real run = K(0.0);
        for (int i = 0; i < n; i++) {
            run += src[j * n + i];
            dst[(j + 1) * stride + i + 1] = run + dst[j * stride + i + 1];
        }
    }
}

/* plane_window (:1518-1525): sum of the inclusive rectangle. */
static real plane_window(const real *p, int n, int i0, int i1, int j0, int j1) {
    int stride = n + 1;
    real a = p[j0 * stride + i0], b = p[j0 * stride + (i1 + 1)];
    real cc = p[(j1 + 1) * stride + i0], d = p[(j1 + 1) * stride + (i1 + 1)];
    return d - b - cc + a;
}

static void window_span(int i, int j, int rad, int n, int *i0, int *i1, int *j0, int *j1) {
    *j0 = j - rad < 0 ? 0 : j - rad;
    *j1 = j + rad > n - 1 ? n - 1 : j + rad;
    *i0 = i - rad < 0 ? 0 : i - rad;
    *i1 = i + rad > n - 1 ? n - 1 : i + rad;
}

/* pond_ruts (:5509-5611): water in a tire cut rises to the mean free surface
 * of the still pond cells within 10 cells; the still cells give the same
 * cubic metres in proportion to their depth above 12 mm. */
real F(oracle_core_pond_ruts)(Core *c) {
    if (c->fine_on) restrict_fine(c);
    const int n = WRES;
    const real area = COARSE * COARSE;
    int any_rut = 0;
    for (int idx = 0; idx < WN; idx++)
        if (R_ISFINITE(c->bed[idx]) && c->bed[idx] < K(-0.012)) {
            any_rut = 1;
            break;
        }
    if (!any_rut) return K(0.0);
    uint8_t *still = (uint8_t *)calloc(WN, 1);
    real *eta_of = zalloc(WN), *mark = zalloc(WN);
    real *count_p = zalloc((size_t)(n + 1) * (n + 1)), *eta_p = zalloc((size_t)(n + 1) * (n + 1));
    real *extra_of = zalloc(WN), *room_of = zalloc(WN);
    for (int idx = 0; idx < WN; idx++) {
        real bed = c->bed[idx], water = c->water[idx];
        if (R_ISFINITE(bed) && R_ISFINITE(water) && bed > K(-0.008) && bed < K(0.03) && water > K(0.018)) {
            still[idx] = 1;
            eta_of[idx] = bed + water;
            mark[idx] = K(1.0);
        }
    }
    prefix_plane(mark, n, count_p);
    prefix_plane(eta_of, n, eta_p);
    real total_need = K(0.0);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            int idx = j * n + i;
            real bed = c->bed[idx], water = c->water[idx];
            if (!R_ISFINITE(bed) || !R_ISFINITE(water) || bed >= K(-0.012)) continue;
            int i0, i1, j0, j1;
            window_span(i, j, 10, n, &i0, &i1, &j0, &j1);
            real count = plane_window(count_p, n, i0, i1, j0, j1);
            real sum = plane_window(eta_p, n, i0, i1, j0, j1);
            if (count < K(1.0)) continue;
            real extra = R_FMAX((sum / count) - bed - water, K(0.0));
            if (extra > K(1.0e-4)) {
                extra_of[idx] = extra;
                total_need += extra * area;
            }
        }
    real moved = K(0.0);
    if (total_need > K(1.0e-8)) {
        real total_room = K(0.0);
        for (int idx = 0; idx < WN; idx++) {
            if (!still[idx]) continue;
            real room = R_FMAX(c->water[idx] - K(0.012), K(0.0));
            if (room > K(1.0e-4)) {
                room_of[idx] = room;
                total_room += room * area;
            }
        }
        if (total_room > K(1.0e-8)) {
            real scale = R_FMIN(total_room / total_need, K(1.0));
            moved = total_need * scale;
            for (int idx = 0; idx < WN; idx++)
                if (extra_of[idx] > K(0.0)) c->water[idx] += extra_of[idx] * scale;
            for (int idx = 0; idx < WN; idx++)
                if (room_of[idx] > K(0.0)) {
                    real give = moved * (room_of[idx] * area) / total_room;
                    c->water[idx] = R_FMAX(c->water[idx] - give / area, K(0.0));
                }
        }
    }
    free(still);
    free(eta_of);
    free(mark);
    free(count_p);
    free(eta_p);
    free(extra_of);
    free(room_of);
    return moved;
}

static int cmp_real(const void *a, const void *b) {
    real x = *(const real *)a, y = *(const real *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

#define SHOVE_RING 4

/* shove_rut_slurry (:5740-5792): slurry scraped off a cut goes, in equal
 * shares, to the non-cut cells of the nearest Chebyshev ring (1 to 4) that
 * has any; a cell that had no slurry takes the cut's grain.  With no such
 * ring the slurry stays on the cut. */
static void shove_rut_slurry(Core *c, const uint8_t *is_rut, const int *sidx, const real *smetres, int ns) {
    const int n = WRES;
    for (int e = 0; e < ns; e++) {
        int idx = sidx[e];
        real metres = smetres[e];
        int i = idx % n, j = idx / n;
        int took = 0;
        for (int d = 1; d <= SHOVE_RING; d++) {
            unsigned count = 0;
            for (int pass = 0; pass < 2; pass++) {
                real share = pass == 0 ? K(0.0) : metres / (real)count;
                int ja = j - d < 0 ? 0 : j - d, jb = j + d > n - 1 ? n - 1 : j + d;
                int ia = i - d < 0 ? 0 : i - d, ib = i + d > n - 1 ? n - 1 : i + d;
                for (int jj = ja; jj <= jb; jj++)
                    for (int ii = ia; ii <= ib; ii++) {
                        int di = ii - i < 0 ? i - ii : ii - i, dj = jj - j < 0 ? j - jj : jj - j;
                        if ((di > dj ? di : dj) != d) continue;
                        int k = jj * n + ii;
                        if (is_rut[k]) continue;
                        if (pass == 0) {
                            count++;
                        } else {
                            if (c->slurry[k] <= K(1.0e-6)) c->grain[k] = c->grain[idx];
                            c->slurry[k] += share;
                        }
                    }
                if (count == 0) break;
            }
            if (count > 0) {
                took = 1;
                break;
            }
        }
        if (!took) c->slurry[idx] += metres;
    }
}

/* seat_rut_water (:5622-5736).  pressed / is_rut: the classify closure
 * sampled per coarse cell (ground = bed + placed, cut flag).  Returns 1 and
 * the pond surface in *eta_out when at least 8 floor cells exist, else 0. */
int F(oracle_core_seat_rut_water)(Core *c, const real *pressed_in, const uint8_t *is_rut, real *eta_out) {
    restrict_fine(c);
    const int n = WRES;
    const real area = COARSE * COARSE;
    real *pressed = zalloc(WN), *etas = zalloc(WN), *floor_src = zalloc(WN);
    real *floor_p = zalloc((size_t)(n + 1) * (n + 1));
    uint8_t *floor_m = (uint8_t *)calloc(WN, 1);
    int *sidx = (int *)malloc(sizeof(int) * WN);
    real *smetres = zalloc(WN);
    int netas = 0, ns = 0, result = 0;
    for (int idx = 0; idx < WN; idx++) {
        real ground = R_ISFINITE(pressed_in[idx]) ? pressed_in[idx] : K(0.0);
        int rut = is_rut[idx] != 0;
        real water = c->water[idx];
        pressed[idx] = ground;
        int on_floor = !rut && ground >= K(-0.02) && ground < K(0.04) &&
                       (water > K(0.008) || c->moisture[idx] > K(0.55));
        floor_m[idx] = (uint8_t)on_floor;
        if (on_floor) etas[netas++] = ground + water;
    }
    if (netas >= 8) {
        /* select_nth_unstable_by at len / 2: the element a full sort puts there. */
        qsort(etas, (size_t)netas, sizeof(real), cmp_real);
        real eta = etas[netas / 2];
        real excess = K(0.0);
        int cut_wet = 0;
        for (int idx = 0; idx < WN; idx++)
            if (is_rut[idx] && c->water[idx] >= K(0.012)) cut_wet = 1;
        if (cut_wet) {
            for (int idx = 0; idx < WN; idx++) floor_src[idx] = floor_m[idx] ? K(1.0) : K(0.0);
            prefix_plane(floor_src, n, floor_p);
            for (int idx = 0; idx < WN; idx++) {
                if (!is_rut[idx] || c->water[idx] < K(0.012)) continue;
                int i = idx % n, j = idx / n, i0, i1, j0, j1;
                window_span(i, j, 8, n, &i0, &i1, &j0, &j1);
                if (plane_window(floor_p, n, i0, i1, j0, j1) <= K(0.0)) continue;
                real limit = R_FMAX(eta - pressed[idx], K(0.0));
                if (c->water[idx] > limit + K(0.004)) {
                    real drop = c->water[idx] - limit;
                    c->water[idx] = limit;
                    excess += drop * area;
                }
                real slurry_room = R_FMAX(eta - pressed[idx] - c->water[idx], K(0.0));
                if (c->slurry[idx] > slurry_room + K(0.004)) {
                    sidx[ns] = idx;
                    smetres[ns] = c->slurry[idx] - slurry_room;
                    ns++;
                    c->slurry[idx] = slurry_room;
                }
            }
        }
        for (int idx = 0; idx < WN; idx++) {
            if (is_rut[idx] || pressed[idx] < K(0.02)) continue;
            real limit = R_FMAX(eta - pressed[idx], K(0.0));
            if (c->water[idx] > limit + K(0.004)) {
                real drop = c->water[idx] - limit;
                c->water[idx] = limit;
                excess += drop * area;
            }
            real slurry_room = R_FMAX(eta - pressed[idx] - c->water[idx], K(0.0));
            if (c->slurry[idx] > slurry_room + K(0.004)) c->slurry[idx] = slurry_room;
        }
        shove_rut_slurry(c, is_rut, sidx, smetres, ns);
        if (excess > K(1.0e-8)) {
            real room = K(0.0);
            for (int idx = 0; idx < WN; idx++)
                if (floor_m[idx]) room += R_FMAX(c->water[idx], K(0.0));
            if (room > K(1.0e-6))
                for (int idx = 0; idx < WN; idx++)
                    if (floor_m[idx] && c->water[idx] > K(0.0))
                        c->water[idx] += excess / area * (c->water[idx] / room);
        }
        *eta_out = eta;
        result = 1;
    }
    free(pressed);
    free(etas);
    free(floor_src);
    free(floor_p);
    free(floor_m);
    free(sidx);
    free(smetres);
    return result;
}

/* The grain plane out (200 x 200 bytes). */
void F(oracle_core_grain)(Core *c, uint8_t *out) { memcpy(out, c->grain, WN); }
