/*
 * oracle/sheet.c -- CPU restatement of the reference's open sheet step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
 * library built from it.  The product path (dirty-gen_amd/) never calls it.
 *
 * It is compiled twice (see oracle/Makefile): REAL=float with
 * -ffp-contract=off restates the reference FP32 arithmetic operation by
 * operation (same order, no fused multiply-add, IEEE division, glibc
 * expf/powf/hypotf like Rust's std on Linux); REAL=double runs the same
 * algorithm in FP64 with the reference's f32 constants widened, so it measures
 * the FP32 rounding error of any implementation.
 *
 * Reference: src/fluid_core.rs of x3-team/Dirty-gen ("silt").  Each function
 * cites the lines it follows.  The open sheet (solid = None, every layer of
 * the plain FluidCore step) and the walled sheet (hydro solids: the creek
 * and ford scenes) are restated.
 *
 * The filter planes are cleared on each box's 2-cell rim when the box's
 * substep starts.  The reference keeps them in a thread-local that outlives
 * boxes and calls; its stale rim values reach only cells at least three cells
 * from every wet cell of the box when the step starts, so a step whose water
 * stays there (one substep, or a slow spread) does not depend on them.
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

/* Reference constants are f32; the FP64 build widens the f32 value. */
#define K(x) ((real)(float)(x))

/*
 * ORACLE_THREADS: the CPU baseline build (Makefile target liboracle_par.so,
 * -O3 -mavx2 -fopenmp, contraction still off).  Every pass that is
 * independent row by row runs in row stripes of at most ORACLE_STRIPES
 * threads, as the reference clamps its rayon stripes (:2154, :2287, :3236,
 * :3652, :4008, clamp(1, 4)), on boxes of at least 48 rows; the results are
 * the serial build's bit for bit.
 */
#ifdef ORACLE_THREADS
#include <omp.h>
#define ORACLE_STRIPES 4
#define ROWS_PAR(b) _Pragma("omp parallel for schedule(static) num_threads(ORACLE_STRIPES) if((b).j1 - (b).j0 >= 47)")
#else
#define ROWS_PAR(b)
#endif

#if defined(REAL_IS_DOUBLE)
#define R_EXP exp
#define R_POW pow
#define R_SQRT sqrt
#define R_HYPOT hypot
#define R_FLOOR floor
#define R_CEIL ceil
#define R_FABS fabs
#define R_FMIN fmin
#define R_FMAX fmax
#define R_ISFINITE(x) isfinite(x)
#else
#define R_EXP expf
#define R_POW powf
#define R_SQRT sqrtf
#define R_HYPOT hypotf
#define R_FLOOR floorf
#define R_CEIL ceilf
#define R_FABS fabsf
#define R_FMIN fminf
#define R_FMAX fmaxf
#define R_ISFINITE(x) isfinite(x)
#endif

static const float GAUSS_F[5] = {0.06f, 0.24f, 0.40f, 0.24f, 0.06f};
static const float STEEP_F[5] = {0.02f, 0.08f, 0.80f, 0.08f, 0.02f};
#define GRAVITY K(9.81)
#define FINE_CELL_M K(0.02)
#define FILM_SPEED_CAP K(1.6)

/* Rust f32::clamp: NaN passes through. */
static inline real clampr(real x, real lo, real hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

static inline size_t clampi(long x, long lo, long hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return (size_t)x;
}

/* ---- constitutive laws: src/fluid_core.rs:150-192 ------------------------ */

/* dynamic_viscosity, src/fluid_core.rs:150-162 */
real F(oracle_dynamic_viscosity)(int grain, real moisture) {
    real x = clampr((clampr(moisture, K(0.0), K(1.0)) - K(0.22)) / K(0.70), K(0.0), K(1.0));
    real thin = x * x * (K(3.0) - K(2.0) * x);
    real stiff, fluid;
    switch (grain) {
    case 1: stiff = K(220.0); fluid = K(4.0); break;   /* Silt */
    case 2: stiff = K(28.0); fluid = K(0.55); break;   /* Sand */
    case 3: stiff = K(0.9); fluid = K(0.9); break;     /* Water */
    case 4: stiff = K(1.5); fluid = K(0.08); break;    /* Other */
    default: stiff = K(420.0); fluid = K(0.35); break; /* Clay */
    }
    return stiff + (fluid - stiff) * thin;
}

/* sheet_mobility_for, src/fluid_core.rs:169-180 */
real F(oracle_sheet_mobility)(real nu) {
    if (nu < K(0.0)) return -nu;
    real boost = K(0.0);
    if (nu < K(12.0)) {
        real t = clampr((K(12.0) - R_FMAX(nu, K(0.0))) / K(12.0), K(0.0), K(1.0));
        boost = t * t * K(1.6);
    }
    return K(0.042) * (K(1.0) + boost);
}

/* sheet_damp, src/fluid_core.rs:184-192 */
real F(oracle_sheet_damp)(real nu) {
    real rel = clampr(R_FMAX(nu, K(0.05)) / K(231.0), K(0.002), K(12.0));
    if (rel >= K(0.42)) return K(1.70) * R_POW(rel, K(0.30));
    real at_silt = K(1.70) * R_POW(K(0.42), K(0.30));
    real t = clampr(rel / K(0.42), K(0.0), K(1.0));
    return at_silt * R_POW(t, K(1.15));
}

/* finite_speed, src/fluid_core.rs:3473-3479 */
static inline real finite_speed(real v, real lim) {
    if (!R_ISFINITE(v)) return K(0.0);
    return clampr(v, -lim, lim);
}

/* wet_sum, src/fluid_core.rs:1305-1319: eight lanes, then the remainder. */
static real wet_sum(const real *h, size_t n, real eps) {
    real lanes[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    size_t full = n / 8 * 8;
    for (size_t c = 0; c < full; c += 8)
        for (int k = 0; k < 8; k++) lanes[k] += (h[c + k] > eps) ? h[c + k] : K(0.0);
    real sum = -K(0.0);
    for (int k = 0; k < 8; k++) sum += lanes[k];
    for (size_t c = full; c < n; c++) sum += (h[c] > eps) ? h[c] : K(0.0);
    return sum;
}

/* ---- work boxes: src/fluid_core.rs:1321-1486, 1527-1657 ----------------- */

typedef struct { size_t i0, i1, j0, j1; } Box;

static int boxes_touch(Box a, Box b) {
    /* saturating +-1 as in boxes_touch, src/fluid_core.rs:1449-1456 */
    size_t ai0 = a.i0 ? a.i0 - 1 : 0, bi0 = b.i0 ? b.i0 - 1 : 0;
    size_t aj0 = a.j0 ? a.j0 - 1 : 0, bj0 = b.j0 ? b.j0 - 1 : 0;
    return ai0 <= b.i1 + 1 && bi0 <= a.i1 + 1 && aj0 <= b.j1 + 1 && bj0 <= a.j1 + 1;
}

/* merge_boxes, src/fluid_core.rs:1458-1486: pairwise merge with swap_remove. */
static void merge_boxes(Box *boxes, size_t *count) {
    for (;;) {
        int did = 0;
        size_t i = 0;
        while (i < *count) {
            size_t j = i + 1;
            while (j < *count) {
                if (boxes_touch(boxes[i], boxes[j])) {
                    Box a = boxes[i], b = boxes[j];
                    boxes[i].i0 = a.i0 < b.i0 ? a.i0 : b.i0;
                    boxes[i].i1 = a.i1 > b.i1 ? a.i1 : b.i1;
                    boxes[i].j0 = a.j0 < b.j0 ? a.j0 : b.j0;
                    boxes[i].j1 = a.j1 > b.j1 ? a.j1 : b.j1;
                    boxes[j] = boxes[*count - 1]; /* swap_remove */
                    (*count)--;
                    did = 1;
                } else {
                    j++;
                }
            }
            i++;
        }
        if (!did) break;
    }
}

static uint32_t uf_root(uint32_t *parent, uint32_t x) {
    while (parent[x] != x) {
        uint32_t up = parent[parent[x]];
        parent[x] = up;
        x = up;
    }
    return x;
}

typedef struct { size_t first; Box bb; } Body;

static int body_cmp(const void *a, const void *b) {
    const Body *x = (const Body *)a, *y = (const Body *)b;
    return (x->first > y->first) - (x->first < y->first);
}

/*
 * sheet_work_boxes, src/fluid_core.rs:1325-1447.  Wet = h > eps.  A solid
 * wet rectangle (or allow_split == 0) is one box with halo 3; otherwise
 * 4-connected wet bodies by row runs, each grown by 6, merged; one merged box
 * or more than 8 fall back to the halo-3 rectangle.  Writes at most `cap`
 * boxes into out, returns the count.
 */
size_t F(oracle_sheet_work_boxes)(const real *h, size_t res, real eps, int allow_split,
                                  size_t *out, size_t cap) {
    size_t i0 = res, i1 = 0, j0 = res, j1 = 0, wet = 0;
    for (size_t j = 0; j < res; j++) {
        const real *row = h + j * res;
        size_t c = 0, first = res, last = 0;
        for (size_t i = 0; i < res; i++)
            if (row[i] > eps) {
                c++;
                if (first == res) first = i;
                last = i;
            }
        if (c == 0) continue;
        wet += c;
        if (first < i0) i0 = first;
        if (last > i1) i1 = last;
        if (j < j0) j0 = j;
        if (j > j1) j1 = j;
    }
    if (wet == 0) return 0;
    Box halo3 = {i0 >= 3 ? i0 - 3 : 0, (i1 + 3 < res - 1) ? i1 + 3 : res - 1,
                 j0 >= 3 ? j0 - 3 : 0, (j1 + 3 < res - 1) ? j1 + 3 : res - 1};
    int solid = wet == (i1 - i0 + 1) * (j1 - j0 + 1);
    Box *result = NULL;
    size_t nres = 0;
    if (!allow_split || solid) {
        result = &halo3;
        nres = 1;
    }
    size_t nruns = 0, runcap = 64;
    size_t (*runs)[2] = NULL;
    uint32_t *parent = NULL;
    Box *bbox = NULL;
    size_t *firsts = NULL;
    Body *bodies = NULL;
    Box *boxes = NULL;
    if (!result) {
        runs = malloc(runcap * sizeof *runs);
        parent = malloc(runcap * sizeof *parent);
        bbox = malloc(runcap * sizeof *bbox);
        firsts = malloc(runcap * sizeof *firsts);
        size_t prev0 = 0, prev1 = 0;
        for (size_t j = j0; j <= j1; j++) {
            const real *row = h + j * res;
            size_t start = nruns;
            size_t i = i0;
            while (i <= i1) {
                if (!(row[i] > eps)) {
                    i++;
                    continue;
                }
                size_t a = i;
                while (i <= i1 && row[i] > eps) i++;
                size_t b = i - 1;
                if (nruns == runcap) {
                    runcap *= 2;
                    runs = realloc(runs, runcap * sizeof *runs);
                    parent = realloc(parent, runcap * sizeof *parent);
                    bbox = realloc(bbox, runcap * sizeof *bbox);
                    firsts = realloc(firsts, runcap * sizeof *firsts);
                }
                uint32_t id = (uint32_t)nruns;
                runs[nruns][0] = a;
                runs[nruns][1] = b;
                parent[nruns] = id;
                bbox[nruns] = (Box){a, b, j, j};
                firsts[nruns] = j * res + a;
                nruns++;
                for (size_t k = prev0; k < prev1; k++) {
                    size_t pa = runs[k][0], pb = runs[k][1];
                    if (pa <= b && a <= pb) {
                        uint32_t ra = uf_root(parent, (uint32_t)k);
                        uint32_t rb = uf_root(parent, id);
                        if (ra != rb) {
                            uint32_t keep = ra < rb ? ra : rb, gone = ra < rb ? rb : ra;
                            parent[gone] = keep;
                            Box g = bbox[gone];
                            Box *kb = &bbox[keep];
                            if (g.i0 < kb->i0) kb->i0 = g.i0;
                            if (g.i1 > kb->i1) kb->i1 = g.i1;
                            if (g.j0 < kb->j0) kb->j0 = g.j0;
                            if (g.j1 > kb->j1) kb->j1 = g.j1;
                            if (firsts[gone] < firsts[keep]) firsts[keep] = firsts[gone];
                        }
                    }
                }
            }
            prev0 = start;
            prev1 = nruns;
        }
        bodies = malloc((nruns ? nruns : 1) * sizeof *bodies);
        size_t nb = 0;
        for (size_t k = 0; k < nruns; k++)
            if (parent[k] == (uint32_t)k) bodies[nb++] = (Body){firsts[k], bbox[k]};
        qsort(bodies, nb, sizeof *bodies, body_cmp);
        boxes = malloc((nb ? nb : 1) * sizeof *boxes);
        for (size_t k = 0; k < nb; k++) {
            Box b = bodies[k].bb;
            boxes[k] = (Box){b.i0 >= 6 ? b.i0 - 6 : 0, (b.i1 + 6 < res - 1) ? b.i1 + 6 : res - 1,
                             b.j0 >= 6 ? b.j0 - 6 : 0, (b.j1 + 6 < res - 1) ? b.j1 + 6 : res - 1};
        }
        size_t nbox = nb;
        merge_boxes(boxes, &nbox);
        if (nbox == 1 || nbox > 8) {
            result = &halo3;
            nres = 1;
        } else {
            result = boxes;
            nres = nbox;
        }
    }
    size_t w = 0;
    for (size_t k = 0; k < nres && w < cap; k++, w++) {
        out[4 * w + 0] = result[k].i0;
        out[4 * w + 1] = result[k].i1;
        out[4 * w + 2] = result[k].j0;
        out[4 * w + 3] = result[k].j1;
    }
    free(runs);
    free(parent);
    free(bbox);
    free(firsts);
    free(bodies);
    free(boxes);
    return nres;
}

/*
 * Flood-fill statement of the same boxes (the reference test's
 * sheet_work_boxes_ref, src/fluid_core.rs:11134-11226).  Kept so the
 * run-length version above can be checked against an independent form.
 */
size_t F(oracle_sheet_work_boxes_flood)(const real *h, size_t res, real eps, int allow_split,
                                        size_t *out, size_t cap) {
    size_t i0 = res, i1 = 0, j0 = res, j1 = 0, wet = 0;
    for (size_t j = 0; j < res; j++)
        for (size_t i = 0; i < res; i++)
            if (h[j * res + i] > eps) {
                wet++;
                if (i < i0) i0 = i;
                if (i > i1) i1 = i;
                if (j < j0) j0 = j;
                if (j > j1) j1 = j;
            }
    if (wet == 0) return 0;
    Box halo3 = {i0 >= 3 ? i0 - 3 : 0, (i1 + 3 < res - 1) ? i1 + 3 : res - 1,
                 j0 >= 3 ? j0 - 3 : 0, (j1 + 3 < res - 1) ? j1 + 3 : res - 1};
    int solid = wet == (i1 - i0 + 1) * (j1 - j0 + 1);
    if (!allow_split || solid) {
        if (cap) {
            out[0] = halo3.i0; out[1] = halo3.i1; out[2] = halo3.j0; out[3] = halo3.j1;
        }
        return 1;
    }
    uint8_t *mark = calloc(res * res, 1);
    size_t *queue = malloc(res * res * sizeof *queue);
    Box *boxes = malloc(res * res * sizeof *boxes);
    size_t nbox = 0;
    for (size_t j = j0; j <= j1; j++)
        for (size_t i = i0; i <= i1; i++) {
            size_t s = j * res + i;
            if (h[s] <= eps || mark[s]) continue;
            size_t ci0 = i, ci1 = i, cj0 = j, cj1 = j, qn = 0;
            mark[s] = 1;
            queue[qn++] = s;
            while (qn) {
                size_t cur = queue[--qn];
                size_t cj = cur / res, ci = cur - cj * res;
                if (ci < ci0) ci0 = ci;
                if (ci > ci1) ci1 = ci;
                if (cj < cj0) cj0 = cj;
                if (cj > cj1) cj1 = cj;
                long nb[4] = {ci > 0 ? (long)cur - 1 : -1, ci + 1 < res ? (long)cur + 1 : -1,
                              cj > 0 ? (long)(cur - res) : -1, cj + 1 < res ? (long)(cur + res) : -1};
                for (int q = 0; q < 4; q++)
                    if (nb[q] >= 0 && !mark[nb[q]] && h[nb[q]] > eps) {
                        mark[nb[q]] = 1;
                        queue[qn++] = (size_t)nb[q];
                    }
            }
            boxes[nbox++] = (Box){ci0 >= 6 ? ci0 - 6 : 0, (ci1 + 6 < res - 1) ? ci1 + 6 : res - 1,
                                  cj0 >= 6 ? cj0 - 6 : 0, (cj1 + 6 < res - 1) ? cj1 + 6 : res - 1};
        }
    merge_boxes(boxes, &nbox);
    size_t n = 0;
    if (nbox == 1 || nbox > 8) {
        if (cap) {
            out[0] = halo3.i0; out[1] = halo3.i1; out[2] = halo3.j0; out[3] = halo3.j1;
        }
        n = 1;
    } else {
        for (size_t k = 0; k < nbox && k < cap; k++) {
            out[4 * k + 0] = boxes[k].i0;
            out[4 * k + 1] = boxes[k].i1;
            out[4 * k + 2] = boxes[k].j0;
            out[4 * k + 3] = boxes[k].j1;
        }
        n = nbox;
    }
    free(mark);
    free(queue);
    free(boxes);
    return n;
}

/*
 * awake_work_boxes, src/fluid_core.rs:1537-1657.  8x8 tiles are awake when a
 * wet cell moves faster than 0.05 m/s or its free surface steps more than
 * 0.02*cell to the +x or +z neighbour; awake tiles grow by one tile and each
 * 4-connected group of grown tiles becomes one box (first tile in raster order).
 */
size_t F(oracle_awake_work_boxes)(const real *h, const real *u, const real *v, const real *bed,
                                  size_t res, real cell, real eps, size_t *out, size_t cap) {
    const size_t T = 8;
    size_t tiles = (res + T - 1) / T, ntile = tiles * tiles;
    uint8_t *hot = calloc(ntile, 1), *grown = calloc(ntile, 1), *seen = calloc(ntile, 1);
    size_t *stack = malloc(ntile * sizeof *stack);
    for (size_t tj = 0; tj < tiles; tj++)
        for (size_t ti = 0; ti < tiles; ti++) {
            size_t i0 = ti * T, j0 = tj * T;
            size_t i1 = i0 + T < res ? i0 + T : res, j1 = j0 + T < res ? j0 + T : res;
            int awake = 0;
            for (size_t j = j0; j < j1 && !awake; j++) {
                size_t row = j * res;
                for (size_t i = i0; i < i1; i++) {
                    size_t idx = row + i;
                    if (h[idx] <= eps) continue;
                    if (R_HYPOT(u[idx], v[idx]) > K(0.05)) {
                        awake = 1;
                        break;
                    }
                    real eta = bed[idx] + h[idx];
                    size_t ip = i + 1 < res ? row + i + 1 : idx;
                    size_t jp = j + 1 < res ? (j + 1) * res + i : idx;
                    real dex = R_FABS(bed[ip] + h[ip] - eta);
                    real dez = R_FABS(bed[jp] + h[jp] - eta);
                    if (R_FMAX(dex, dez) > K(0.02) * cell) {
                        awake = 1;
                        break;
                    }
                }
            }
            if (awake) hot[tj * tiles + ti] = 1;
        }
    for (size_t tj = 0; tj < tiles; tj++)
        for (size_t ti = 0; ti < tiles; ti++) {
            if (!hot[tj * tiles + ti]) continue;
            for (int dj = -1; dj <= 1; dj++)
                for (int di = -1; di <= 1; di++) {
                    long ni = (long)ti + di, nj = (long)tj + dj;
                    if (ni >= 0 && nj >= 0 && ni < (long)tiles && nj < (long)tiles)
                        grown[nj * tiles + ni] = 1;
                }
        }
    size_t nbox = 0;
    for (size_t s = 0; s < ntile; s++) {
        if (!grown[s] || seen[s]) continue;
        seen[s] = 1;
        size_t sp = 0;
        stack[sp++] = s;
        size_t ti0 = s % tiles, ti1 = ti0, tj0 = s / tiles, tj1 = tj0;
        while (sp) {
            size_t cur = stack[--sp];
            size_t tj = cur / tiles, ti = cur - tj * tiles;
            if (ti < ti0) ti0 = ti;
            if (ti > ti1) ti1 = ti;
            if (tj < tj0) tj0 = tj;
            if (tj > tj1) tj1 = tj;
            size_t nb[4];
            int nn = 0;
            if (ti > 0) nb[nn++] = tj * tiles + ti - 1;
            if (ti + 1 < tiles) nb[nn++] = tj * tiles + ti + 1;
            if (tj > 0) nb[nn++] = (tj - 1) * tiles + ti;
            if (tj + 1 < tiles) nb[nn++] = (tj + 1) * tiles + ti;
            for (int q = 0; q < nn; q++)
                if (grown[nb[q]] && !seen[nb[q]]) {
                    seen[nb[q]] = 1;
                    stack[sp++] = nb[q];
                }
        }
        if (nbox < cap) {
            size_t e_i = (ti1 + 1) * T < res ? (ti1 + 1) * T : res;
            size_t e_j = (tj1 + 1) * T < res ? (tj1 + 1) * T : res;
            out[4 * nbox + 0] = ti0 * T;
            out[4 * nbox + 1] = e_i - 1;
            out[4 * nbox + 2] = tj0 * T;
            out[4 * nbox + 3] = e_j - 1;
        }
        nbox++;
    }
    free(hot);
    free(grown);
    free(seen);
    free(stack);
    return nbox;
}

/* ---- stencil helpers: src/fluid_core.rs:3489-3539 ----------------------- */

static inline real vorticity_at(const real *u, const real *v, size_t res, size_t i, size_t j,
                                real cell) {
    size_t ip = i + 1 < res ? i + 1 : res - 1, im = i ? i - 1 : 0;
    size_t jp = j + 1 < res ? j + 1 : res - 1, jm = j ? j - 1 : 0;
    real dvdx = (v[j * res + ip] - v[j * res + im]) / (K(2.0) * cell);
    real dudz = (u[jp * res + i] - u[jm * res + i]) / (K(2.0) * cell);
    return dvdx - dudz;
}

static inline void grad(const real *p, size_t res, size_t i, size_t j, real cell, real *dx,
                        real *dz) {
    size_t ip = i + 1 < res ? i + 1 : res - 1, im = i ? i - 1 : 0;
    size_t jp = j + 1 < res ? j + 1 : res - 1, jm = j ? j - 1 : 0;
    *dx = (p[j * res + ip] - p[j * res + im]) / (K(2.0) * cell);
    *dz = (p[jp * res + i] - p[jm * res + i]) / (K(2.0) * cell);
}

static inline real lap(const real *p, size_t res, size_t i, size_t j, real cell) {
    size_t idx = j * res + i;
    size_t ip = i + 1 < res ? i + 1 : res - 1, im = i ? i - 1 : 0;
    size_t jp = j + 1 < res ? j + 1 : res - 1, jm = j ? j - 1 : 0;
    real s = p[j * res + ip] + p[j * res + im] + p[jp * res + i] + p[jm * res + i] - K(4.0) * p[idx];
    return s / (cell * cell);
}

/*
 * One vorticity-confinement kick at (i, j) (confine_one / confine_interior /
 * confine_jacobi, src/fluid_core.rs:2402-2503, 3442-3471).  The vorticity is
 * read from (su, sv); the kick is added to (u, v).  Gauss-Seidel passes the
 * live planes as both; Jacobi passes a snapshot as (su, sv).  Clamped indices
 * equal the interior stencil away from the grid edge.
 */
static inline void confine(const real *su, const real *sv, real *u, real *v, size_t res, size_t i,
                           size_t j, real cell, real dt, real eps) {
    real w = vorticity_at(su, sv, res, i, j, cell);
    if (!R_ISFINITE(w) || R_FABS(w) < K(1.0e-4)) return;
    real wxp = R_FABS(vorticity_at(su, sv, res, i + 1 < res ? i + 1 : res - 1, j, cell));
    real wxm = R_FABS(vorticity_at(su, sv, res, i ? i - 1 : 0, j, cell));
    real wzp = R_FABS(vorticity_at(su, sv, res, i, j + 1 < res ? j + 1 : res - 1, cell));
    real wzm = R_FABS(vorticity_at(su, sv, res, i, j ? j - 1 : 0, cell));
    real nx = (wxp - wxm) / (K(2.0) * cell);
    real nz = (wzp - wzm) / (K(2.0) * cell);
    real du = eps * dt * (-nz * w);
    real dv = eps * dt * (nx * w);
    if (R_ISFINITE(du) && R_ISFINITE(dv) && du * du + dv * dv <= K(0.0625)) {
        size_t idx = j * res + i;
        u[idx] += du;
        v[idx] += dv;
    }
}

/* accum5, src/fluid_core.rs:3681-3688: left to right, first product unadded. */
static inline real accum5(real a0, real a1, real a2, real a3, real a4, const float *k) {
    real a = a0 * K(k[0]);
    a += a1 * K(k[1]);
    a += a2 * K(k[2]);
    a += a3 * K(k[3]);
    a += a4 * K(k[4]);
    return a;
}

/*
 * blur_box, src/fluid_core.rs:3541-3569 with blur_axis_x/z (:3700-3757): a
 * 5-tap pass along x on the box rows into tmp, then along z into dst.  Taps
 * clamp at the grid edge, where the sum starts from 0.0 (accum5_clamped).
 */
static void blur_box(const real *src, real *dst, real *tmp, size_t res, Box b, const float *k) {
    if (b.i0 > b.i1 || b.j0 > b.j1 || res == 0) return;
    ROWS_PAR(b)
    for (size_t j = b.j0; j <= b.j1; j++) {
        const real *row = src + j * res;
        for (size_t i = b.i0; i <= b.i1; i++) {
            real a;
            if (i >= 2 && i + 2 < res) {
                a = accum5(row[i - 2], row[i - 1], row[i], row[i + 1], row[i + 2], k);
            } else {
                a = K(0.0);
                for (int t = -2; t <= 2; t++) a += row[clampi((long)i + t, 0, (long)res - 1)] * K(k[t + 2]);
            }
            tmp[j * res + i] = a;
        }
    }
    ROWS_PAR(b)
    for (size_t j = b.j0; j <= b.j1; j++) {
        for (size_t i = b.i0; i <= b.i1; i++) {
            real a;
            if (j >= 2 && j + 2 < res) {
                a = accum5(tmp[(j - 2) * res + i], tmp[(j - 1) * res + i], tmp[j * res + i],
                           tmp[(j + 1) * res + i], tmp[(j + 2) * res + i], k);
            } else {
                a = K(0.0);
                for (int t = -2; t <= 2; t++)
                    a += tmp[clampi((long)j + t, 0, (long)res - 1) * res + i] * K(k[t + 2]);
            }
            dst[j * res + i] = a;
        }
    }
}

/* ---- open transport: src/fluid_core.rs:2847-3215 ------------------------ */

/* dry_bank, src/fluid_core.rs:2843-2845 */
static inline int dry_bank(const real *h, const real *bed, size_t donor, size_t recv) {
    return h[recv] <= K(1.0e-6) && bed[recv] > bed[donor] + h[donor] + K(1.0e-3);
}

/* flux_open, src/fluid_core.rs:3412-3440 */
static inline void flux_open(const real *h, const uint8_t *gate, size_t a, size_t b, real vel,
                             real cell, real dt, real frac_cap, real *removed, real *added) {
    size_t donor = vel > K(0.0) ? a : b, recv = vel > K(0.0) ? b : a;
    real room = h[donor] - removed[donor];
    if (room <= K(1.0e-8)) return;
    if (gate[recv] == 0 && h[recv] <= K(1.0e-6)) return;
    real frac = R_FMIN(R_FABS(vel) * dt / cell, frac_cap);
    real dh = R_FMIN(h[donor] * frac, room);
    if (dh <= K(1.0e-8)) return;
    removed[donor] += dh;
    added[recv] += dh;
}

static void apply_moves(real *h, const real *removed, const real *added, size_t res, Box b) {
    ROWS_PAR(b)
    for (size_t j = b.j0; j <= b.j1; j++)
        for (size_t i = b.i0; i <= b.i1; i++) {
            size_t idx = j * res + i;
            real r = removed[idx], a = added[idx];
            real moved = R_FMAX(h[idx] - r + a, K(0.0));
            h[idx] = (r > K(0.0) || a > K(0.0)) ? moved : h[idx];
        }
}

/*
 * Serial edge loop (transport_open_ref, src/fluid_core.rs:3139-3215; the
 * rayon form transport_open_par :3219-3397 computes the same edges).  Edges
 * interior to the box only: its outer faces act as walls.
 */
void F(oracle_transport_open_ref)(real *h, const real *vel, const uint8_t *gate, size_t res,
                                  real cell, real dt, size_t i0, size_t i1, size_t j0, size_t j1,
                                  int along_x, const real *bed, real frac_cap, real *removed,
                                  real *added) {
    Box b = {i0, i1, j0, j1};
    for (size_t j = j0; j <= j1; j++)
        for (size_t i = i0; i <= i1; i++) removed[j * res + i] = added[j * res + i] = K(0.0);
    if (along_x) {
        for (size_t j = j0; j <= j1; j++)
            for (size_t i = i0; i < i1; i++) {
                size_t a = j * res + i, bb = a + 1;
                real speed = K(0.5) * (vel[a] + vel[bb]);
                if (R_FABS(speed) < K(1.0e-5)) continue;
                if (bed) {
                    size_t d = speed > K(0.0) ? a : bb, r = speed > K(0.0) ? bb : a;
                    if (dry_bank(h, bed, d, r)) continue;
                }
                flux_open(h, gate, a, bb, speed, cell, dt, frac_cap, removed, added);
            }
    } else {
        for (size_t j = j0; j < j1; j++)
            for (size_t i = i0; i <= i1; i++) {
                size_t a = j * res + i, bb = a + res;
                real speed = K(0.5) * (vel[a] + vel[bb]);
                if (R_FABS(speed) < K(1.0e-5)) continue;
                if (bed) {
                    size_t d = speed > K(0.0) ? a : bb, r = speed > K(0.0) ? bb : a;
                    if (dry_bank(h, bed, d, r)) continue;
                }
                flux_open(h, gate, a, bb, speed, cell, dt, frac_cap, removed, added);
            }
    }
    apply_moves(h, removed, added, res, b);
}

/* open_edge_dh, src/fluid_core.rs:2887-2911 */
static inline real open_edge_dh(real hd, real hr, real remd, uint8_t gate_r, int use_bed, real bed_d,
                                real bed_r, real speed, real dt, real cell, real frac_cap) {
    int live = !(R_FABS(speed) < K(1.0e-5));
    int bank = use_bed && hr <= K(1.0e-6) && bed_r > bed_d + hd + K(1.0e-3);
    real room = hd - remd;
    int gate_shut = gate_r == 0 && hr <= K(1.0e-6);
    real frac = R_FMIN(R_FABS(speed) * dt / cell, frac_cap);
    real dh = R_FMIN(hd * frac, room);
    int ok = live && !bank && !(room <= K(1.0e-8)) && !gate_shut && !(dh <= K(1.0e-8));
    return ok ? dh : K(0.0);
}

/*
 * Branch-free row form (transport_open_rows_body, src/fluid_core.rs:2998-3137):
 * left-going fluxes first, then right-going, then each cell sums its two
 * edges in the serial order.  Equal to the serial loop bit for bit.
 */
#ifdef ORACLE_THREADS
/*
 * The z pass of transport_open_rows in row stripes.  The serial loop walks
 * the faces (j, j+1) in order: the up flux of a face (donor j+1) needs no
 * room, the down flux (donor j) has room h_j less the up flux j already gave
 * through face (j-1, j), and each row's removed and added sums collect the
 * two faces' fluxes in that order.  Every face's up flux first, then every
 * down flux, then each row's sums in the serial order: the same values.
 */
static void transport_z_rows_par(real *h, const real *vel, const uint8_t *gate, size_t res, real cell, real dt,
                                 Box b, const real *bed, real frac_cap, real *removed, real *added) {
    size_t m = b.i1 + 1 - b.i0, nf = b.j1 - b.j0;
    int ub = bed != NULL;
    static _Thread_local real *up_down = NULL;
    static _Thread_local size_t cap = 0;
    if (cap < 2 * nf * m) {
        free(up_down);
        up_down = malloc(2 * nf * m * sizeof(real));
        cap = 2 * nf * m;
    }
    real *up = up_down, *down = up_down + nf * m;
    ROWS_PAR(b)
    for (size_t f = 0; f < nf; f++) {
        size_t lo = (b.j0 + f) * res + b.i0;
        for (size_t i = 0; i < m; i++) {
            size_t a = lo + i, c = lo + res + i;
            real speed = K(0.5) * (vel[a] + vel[c]);
            real dh = open_edge_dh(h[c], h[a], K(0.0), gate[a], ub, ub ? bed[c] : 0, ub ? bed[a] : 0, speed, dt, cell,
                                   frac_cap);
            up[f * m + i] = speed > K(0.0) ? K(0.0) : dh;
        }
    }
    ROWS_PAR(b)
    for (size_t f = 0; f < nf; f++) {
        size_t lo = (b.j0 + f) * res + b.i0;
        for (size_t i = 0; i < m; i++) {
            size_t a = lo + i, c = lo + res + i;
            real speed = K(0.5) * (vel[a] + vel[c]);
            real remd = f > 0 ? K(0.0) + up[(f - 1) * m + i] : K(0.0);
            real dh = open_edge_dh(h[a], h[c], remd, gate[c], ub, ub ? bed[a] : 0, ub ? bed[c] : 0, speed, dt, cell,
                                   frac_cap);
            down[f * m + i] = speed > K(0.0) ? dh : K(0.0);
        }
    }
    ROWS_PAR(b)
    for (size_t j = b.j0; j <= b.j1; j++) {
        size_t f = j - b.j0, lo = j * res + b.i0;
        for (size_t i = 0; i < m; i++) {
            real r = f > 0 ? K(0.0) + up[(f - 1) * m + i] : K(0.0);
            real a = f > 0 ? K(0.0) + down[(f - 1) * m + i] : K(0.0);
            if (j < b.j1) {
                r = r + down[f * m + i];
                a = a + up[f * m + i];
            }
            removed[lo + i] = r;
            added[lo + i] = a;
        }
    }
    apply_moves(h, removed, added, res, b);
}
#endif

void F(oracle_transport_open_rows)(real *h, const real *vel, const uint8_t *gate, size_t res,
                                   real cell, real dt, size_t i0, size_t i1, size_t j0, size_t j1,
                                   int along_x, const real *bed, real frac_cap, real *removed,
                                   real *added) {
    size_t m = i1 + 1 - i0;
    int ub = bed != NULL;
#ifdef ORACLE_THREADS
    Box bt = {i0, i1, j0, j1};
    (void)bt;
    if (!along_x) {
        transport_z_rows_par(h, vel, gate, res, cell, dt, bt, bed, frac_cap, removed, added);
        return;
    }
#endif
    if (along_x) {
#ifdef ORACLE_THREADS
#pragma omp parallel num_threads(ORACLE_STRIPES) if (j1 - j0 >= 47)
        {
#endif
        real *fl = calloc(m, sizeof *fl), *fr = calloc(m, sizeof *fr);
#ifdef ORACLE_THREADS
#pragma omp for schedule(static)
#endif
        for (size_t j = j0; j <= j1; j++) {
            size_t lo = j * res + i0;
            real *hr = h + lo;
            const real *vr = vel + lo;
            const uint8_t *gr = gate + lo;
            const real *br = ub ? bed + lo : NULL;
            size_t edges = m - 1;
            for (size_t k = 0; k < edges; k++) {
                real speed = K(0.5) * (vr[k] + vr[k + 1]);
                real dh = open_edge_dh(hr[k + 1], hr[k], K(0.0), gr[k], ub, ub ? br[k + 1] : 0,
                                       ub ? br[k] : 0, speed, dt, cell, frac_cap);
                fl[k] = speed > K(0.0) ? K(0.0) : dh;
            }
            for (size_t k = 0; k < edges; k++) {
                real speed = K(0.5) * (vr[k] + vr[k + 1]);
                real prev = k > 0 ? fl[k - 1] : K(0.0);
                real dh = open_edge_dh(hr[k], hr[k + 1], K(0.0) + prev, gr[k + 1], ub,
                                       ub ? br[k] : 0, ub ? br[k + 1] : 0, speed, dt, cell, frac_cap);
                fr[k] = speed > K(0.0) ? dh : K(0.0);
            }
            fl[edges] = K(0.0);
            fr[edges] = K(0.0);
            real *rem = removed + lo, *add = added + lo;
            rem[0] = (K(0.0) + K(0.0)) + fr[0];
            add[0] = (K(0.0) + K(0.0)) + fl[0];
            for (size_t c = 1; c < m; c++) {
                rem[c] = (K(0.0) + fl[c - 1]) + fr[c];
                add[c] = (K(0.0) + fr[c - 1]) + fl[c];
            }
            for (size_t c = 0; c < m; c++) {
                real moved = R_FMAX(hr[c] - rem[c] + add[c], K(0.0));
                hr[c] = (rem[c] > K(0.0) || add[c] > K(0.0)) ? moved : hr[c];
            }
        }
        free(fl);
        free(fr);
#ifdef ORACLE_THREADS
        }
#endif
        return;
    }
    for (size_t j = j0; j <= j1; j++)
        for (size_t c = 0; c < m; c++) removed[j * res + i0 + c] = added[j * res + i0 + c] = K(0.0);
    for (size_t j = j0; j < j1; j++) {
        size_t lo = j * res + i0;
        const real *ha = h + lo, *hb = h + lo + res, *va = vel + lo, *vb = vel + lo + res;
        const uint8_t *ga = gate + lo, *gb = gate + lo + res;
        const real *ba = ub ? bed + lo : NULL, *bb = ub ? bed + lo + res : NULL;
        real *ra = removed + lo, *rb = removed + lo + res, *aa = added + lo, *ab = added + lo + res;
        for (size_t i = 0; i < m; i++) {
            real speed = K(0.5) * (va[i] + vb[i]);
            real dh = open_edge_dh(hb[i], ha[i], K(0.0), ga[i], ub, ub ? bb[i] : 0, ub ? ba[i] : 0,
                                   speed, dt, cell, frac_cap);
            real up = speed > K(0.0) ? K(0.0) : dh;
            rb[i] += up;
            aa[i] += up;
        }
        for (size_t i = 0; i < m; i++) {
            real speed = K(0.5) * (va[i] + vb[i]);
            real dh = open_edge_dh(ha[i], hb[i], ra[i], gb[i], ub, ub ? ba[i] : 0, ub ? bb[i] : 0,
                                   speed, dt, cell, frac_cap);
            real down = speed > K(0.0) ? dh : K(0.0);
            ra[i] += down;
            ab[i] += down;
        }
    }
    Box b = {i0, i1, j0, j1};
    apply_moves(h, removed, added, res, b);
}

/* ---- the substep: src/fluid_core.rs:1879-2082 (solid = None) ------------- */

typedef struct {
    int wind;          /* 0 Gravity, 1 Wind */
    real wx, wz;       /* wind velocity */
    real vorticity;
} Drive;

typedef struct {
    real *a, *b, *c, *d, *e, *removed, *added, *su, *sv;
} Scratch;

/* Confinement schedule: the reference's own choice, or Jacobi everywhere. */
enum { CONFINE_REFERENCE = 0, CONFINE_JACOBI = 1, CONFINE_GAUSS_SEIDEL = 2 };

/*
 * Gravity (or wind blend) on the wet cells of the box, then one confinement
 * sweep.  The reference sweeps in place (forward Gauss-Seidel) for a small
 * box or a coarse sheet stepped beside the other layer, and from a snapshot
 * (Jacobi) on the 2 cm sheet and on boxes of at least 48 x 32 when row
 * splits are allowed: force_gravity_open :2204-2261, force_gravity_par
 * :2263-2378, gravity_jacobi_serial :2533-2595, force_wind_open :2633-2682.
 */
static void force_open(const real *h, real *u, real *v, const real *nu, const real *p, size_t res,
                       real cell, real dt, Drive drive, real eps, Box b, int inner_par,
                       int confine_mode, Scratch *sc) {
    int jacobi;
    if (confine_mode == CONFINE_JACOBI) {
        jacobi = 1;
    } else if (confine_mode == CONFINE_GAUSS_SEIDEL) {
        jacobi = 0;
    } else {
        int big = (b.j1 - b.j0 >= 47) && (b.i1 - b.i0 >= 31) && inner_par;
        int fine = !drive.wind && cell <= FINE_CELL_M + K(1.0e-6);
        jacobi = fine || big;
    }
    real two = K(2.0) * cell;
    real blend = K(1.0) - R_EXP(-dt * K(1.6));
    if (!jacobi) {
        for (size_t j = b.j0; j <= b.j1; j++)
            for (size_t i = b.i0; i <= b.i1; i++) {
                size_t idx = j * res + i;
                if (h[idx] <= eps) continue;
                if (drive.wind) {
                    u[idx] += (drive.wx - u[idx]) * blend;
                    v[idx] += (drive.wz - v[idx]) * blend;
                } else {
                    real px, pz;
                    if (i >= 2 && j >= 2 && i + 2 < res && j + 2 < res) {
                        px = (p[idx + 1] - p[idx - 1]) / two;
                        pz = (p[idx + res] - p[idx - res]) / two;
                    } else {
                        grad(p, res, i, j, cell, &px, &pz);
                    }
                    real mob = F(oracle_sheet_mobility)(nu[idx]);
                    u[idx] += -GRAVITY * px * mob * dt;
                    v[idx] += -GRAVITY * pz * mob * dt;
                }
                if (drive.vorticity > K(0.0)) confine(u, v, u, v, res, i, j, cell, dt, drive.vorticity);
            }
        return;
    }
    ROWS_PAR(b)
    for (size_t j = b.j0; j <= b.j1; j++)
        for (size_t i = b.i0; i <= b.i1; i++) {
            size_t idx = j * res + i;
            if (h[idx] <= eps) continue;
            if (drive.wind) {
                u[idx] += (drive.wx - u[idx]) * blend;
                v[idx] += (drive.wz - v[idx]) * blend;
            } else {
                real px, pz;
                if (i >= 2 && j >= 2 && i + 2 < res && j + 2 < res) {
                    px = (p[idx + 1] - p[idx - 1]) / two;
                    pz = (p[idx + res] - p[idx - res]) / two;
                } else {
                    grad(p, res, i, j, cell, &px, &pz);
                }
                real mob = F(oracle_sheet_mobility)(nu[idx]);
                u[idx] += -GRAVITY * px * mob * dt;
                v[idx] += -GRAVITY * pz * mob * dt;
            }
        }
    if (drive.vorticity <= K(0.0)) return;
    memcpy(sc->su, u, res * res * sizeof(real));
    memcpy(sc->sv, v, res * res * sizeof(real));
    ROWS_PAR(b)
    for (size_t j = b.j0; j <= b.j1; j++)
        for (size_t i = b.i0; i <= b.i1; i++) {
            if (h[j * res + i] <= eps) continue;
            confine(sc->su, sc->sv, u, v, res, i, j, cell, dt, drive.vorticity);
        }
}

static void transport_rows(real *h, const real *vel, const uint8_t *gate, size_t res, real cell,
                           real dt, Box b, int along_x, const real *bed, real frac_cap, Scratch *sc) {
    F(oracle_transport_open_rows)(h, vel, gate, res, cell, dt, b.i0, b.i1, b.j0, b.j1, along_x, bed,
                                  frac_cap, sc->removed, sc->added);
}

/*
 * The substep's surface and pressure planes on box b (:1903-1935): a = eta =
 * bed + h on the box and a 2-cell rim, b = G5(eta), c = S5(eta),
 * d = 0.45 (0.62 b + 0.38 c) + 0.55 eta.  Each box starts from cleared filter
 * planes on its rim (see the scratch note at the top of this file).
 */
static void surface_pressure(const real *h, const real *bed, size_t res, Box b, int wind_open, Scratch *sc) {
    size_t hi0 = b.i0 >= 2 ? b.i0 - 2 : 0, hi1 = (b.i1 + 2 < res - 1) ? b.i1 + 2 : res - 1;
    size_t hj0 = b.j0 >= 2 ? b.j0 - 2 : 0, hj1 = (b.j1 + 2 < res - 1) ? b.j1 + 2 : res - 1;
    ROWS_PAR(b)
    for (size_t j = hj0; j <= hj1; j++)
        for (size_t i = hi0; i <= hi1; i++) {
            size_t k = j * res + i;
            sc->b[k] = sc->c[k] = sc->d[k] = sc->e[k] = K(0.0);
            sc->a[k] = bed[k] + h[k];
        }
    if (!wind_open) {
        blur_box(sc->a, sc->b, sc->e, res, b, GAUSS_F);
        blur_box(sc->a, sc->c, sc->e, res, b, STEEP_F);
        ROWS_PAR(b)
        for (size_t j = b.j0; j <= b.j1; j++)
            for (size_t i = b.i0; i <= b.i1; i++) {
                size_t idx = j * res + i;
                real filtered = K(0.62) * sc->b[idx] + K(0.38) * sc->c[idx];
                sc->d[idx] = K(0.45) * filtered + K(0.55) * sc->a[idx];
            }
    }
}

static void substep(real *h, real *u, real *v, const real *bed, const real *nu, const uint8_t *gate,
                    size_t res, real cell, real dt, Drive drive, real eps, Box b, int banks,
                    int do_flux, int inner_par, int confine_mode, Scratch *sc) {
    surface_pressure(h, bed, res, b, drive.wind, sc);
    force_open(h, u, v, nu, sc->d, res, cell, dt, drive, eps, b, inner_par, confine_mode, sc);
    int project = !drive.wind && cell > FINE_CELL_M + K(1.0e-6);
    if (project) {
        ROWS_PAR(b)
        for (size_t j = b.j0; j <= b.j1; j++)
            for (size_t i = b.i0; i <= b.i1; i++) {
                size_t idx = j * res + i;
                size_t ip = i + 1 < res ? i + 1 : res - 1, im = i ? i - 1 : 0;
                size_t jp = j + 1 < res ? j + 1 : res - 1, jm = j ? j - 1 : 0;
                real ux = (u[j * res + ip] - u[j * res + im]) / (K(2.0) * cell);
                real vz = (v[jp * res + i] - v[jm * res + i]) / (K(2.0) * cell);
                real div = ux + vz;
                real lap_h = lap(sc->a, res, i, j, cell);
                real overload = K(1.0) + clampr(nu[idx] / K(231.0), K(0.0), K(6.0));
                sc->b[idx] = div * overload - K(0.015) * overload * lap_h;
            }
        blur_box(sc->b, sc->c, sc->e, res, b, GAUSS_F);
        blur_box(sc->b, sc->a, sc->e, res, b, STEEP_F);
        ROWS_PAR(b)
        for (size_t j = b.j0; j <= b.j1; j++)
            for (size_t i = b.i0; i <= b.i1; i++) {
                size_t idx = j * res + i;
                sc->d[idx] = K(0.62) * sc->c[idx] + K(0.38) * sc->a[idx];
            }
    }
    /* Damping, memoised per viscosity bit pattern: :2014-2042 (the memo is a
     * cache of a pure function, so one per row gives the same values). */
    real lim = drive.wind ? R_FMAX(R_HYPOT(drive.wx, drive.wz), FILM_SPEED_CAP) : FILM_SPEED_CAP;
    ROWS_PAR(b)
    for (size_t j = b.j0; j <= b.j1; j++) {
        int memo_set = 0;
        real memo_nu = K(0.0), memo_decay = K(1.0);
        for (size_t i = b.i0; i <= b.i1; i++) {
            size_t idx = j * res + i;
            if (h[idx] <= eps) continue;
            if (project) {
                real qx, qz;
                grad(sc->d, res, i, j, cell, &qx, &qz);
                u[idx] += K(0.012) * dt * qx;
                v[idx] += K(0.012) * dt * qz;
            }
            real nu_here = nu[idx];
            if (!memo_set || memcmp(&nu_here, &memo_nu, sizeof(real)) != 0) {
                memo_set = 1;
                memo_nu = nu_here;
                memo_decay = R_EXP(-F(oracle_sheet_damp)(nu_here) * dt);
            }
            u[idx] = finite_speed(u[idx] * memo_decay, lim);
            v[idx] = finite_speed(v[idx] * memo_decay, lim);
        }
    }
    if (!do_flux) return;
    const real *open_bed = banks ? bed : NULL;
    transport_rows(h, u, gate, res, cell, dt, b, 1, open_bed, K(0.45), sc);
    transport_rows(h, v, gate, res, cell, dt, b, 0, open_bed, K(0.45), sc);
    ROWS_PAR(b)
    for (size_t j = b.j0; j <= b.j1; j++)
        for (size_t i = b.i0; i <= b.i1; i++) {
            size_t idx = j * res + i;
            if (eps > K(0.0) && h[idx] < eps) {
                h[idx] = K(0.0);
                u[idx] = K(0.0);
                v[idx] = K(0.0);
            }
        }
}

/* pace3_slices, src/fluid_core.rs:1835-1868 */
static void pace3_slices(real *h, real *u, real *v, const uint8_t *gate, size_t res, real cell,
                         real dts, uint32_t nsub, real eps, Box b, Scratch *sc) {
    for (uint32_t s = 0; s < nsub; s++) {
        transport_rows(h, u, gate, res, cell, dts, b, 1, NULL, K(0.45), sc);
        transport_rows(h, v, gate, res, cell, dts, b, 0, NULL, K(0.45), sc);
    }
    if (eps > K(0.0))
        for (size_t j = b.j0; j <= b.j1; j++)
            for (size_t i = b.i0; i <= b.i1; i++) {
                size_t idx = j * res + i;
                if (h[idx] < eps) h[idx] = u[idx] = v[idx] = K(0.0);
            }
}

/* ---- walled sheet (hydro solids): src/fluid_core.rs:1879-2082 with
 * solid = Some, 2123-2200, 3894-4563 -------------------------------------- */

/* A 4-neighbour offset table in the reference's order: (-1,0) (1,0) (0,-1) (0,1). */
static const int NB4[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};

static inline int in_grid(long i, long j, size_t res) {
    return i >= 0 && j >= 0 && i < (long)res && j < (long)res;
}

/* solid_near, :4565-4582 */
static int solid_near(const uint8_t *mask, size_t res, size_t i, size_t j, int radius) {
    for (int dj = -radius; dj <= radius; dj++)
        for (int di = -radius; di <= radius; di++) {
            if (di == 0 && dj == 0) continue;
            long ni = (long)i + di, nj = (long)j + dj;
            if (!in_grid(ni, nj, res)) continue;
            if (mask[(size_t)nj * res + (size_t)ni]) return 1;
        }
    return 0;
}

/* lift_solid_surface, :4225-4276: each solid box cell takes the mean eta of
 * the wet open cells within two cells (tmp holds the means first). */
void F(oracle_lift_solid_surface)(real *eta, const real *h, const uint8_t *solid, size_t res, size_t i0,
                                  size_t i1, size_t j0, size_t j1, real eps, real *tmp) {
    for (size_t j = j0; j <= j1; j++)
        for (size_t i = i0; i <= i1; i++) {
            size_t idx = j * res + i;
            if (!solid[idx]) continue;
            real acc = K(0.0), n = K(0.0);
            for (int dj = -2; dj <= 2; dj++)
                for (int di = -2; di <= 2; di++) {
                    if (di == 0 && dj == 0) continue;
                    long ni = (long)i + di, nj = (long)j + dj;
                    if (!in_grid(ni, nj, res)) continue;
                    size_t nidx = (size_t)nj * res + (size_t)ni;
                    if (solid[nidx] || h[nidx] <= eps) continue;
                    acc += eta[nidx];
                    n += K(1.0);
                }
            tmp[idx] = n > K(0.0) ? acc / n : eta[idx];
        }
    for (size_t j = j0; j <= j1; j++)
        for (size_t i = i0; i <= i1; i++)
            if (solid[j * res + i]) eta[j * res + i] = tmp[j * res + i];
}

/* write_noslip_ghost, :4278-4326: a solid box cell takes minus the mean
 * velocity of its wet open 4-neighbours (0 without one), from the values
 * before any ghost is written. */
void F(oracle_noslip_ghost)(real *u, real *v, const real *h, const uint8_t *solid, size_t res, size_t i0,
                            size_t i1, size_t j0, size_t j1, real eps, real *gu, real *gv) {
    for (size_t j = j0; j <= j1; j++)
        for (size_t i = i0; i <= i1; i++) {
            size_t idx = j * res + i;
            if (!solid[idx]) continue;
            real au = K(0.0), av = K(0.0), n = K(0.0);
            for (int k = 0; k < 4; k++) {
                long ni = (long)i + NB4[k][0], nj = (long)j + NB4[k][1];
                if (!in_grid(ni, nj, res)) continue;
                size_t nidx = (size_t)nj * res + (size_t)ni;
                if (solid[nidx] || h[nidx] <= eps) continue;
                au += u[nidx];
                av += v[nidx];
                n += K(1.0);
            }
            gu[idx] = n > K(0.0) ? -au / n : K(0.0);
            gv[idx] = n > K(0.0) ? -av / n : K(0.0);
        }
    for (size_t j = j0; j <= j1; j++)
        for (size_t i = i0; i <= i1; i++) {
            size_t idx = j * res + i;
            if (!solid[idx]) continue;
            u[idx] = gu[idx];
            v[idx] = gv[idx];
        }
}

/* block_into_solid, :4328-4380: the into-wall part of a wet open cell's
 * velocity is dropped, wall by wall in neighbour order. */
void F(oracle_block_into_solid)(real *u, real *v, const real *h, const uint8_t *solid, size_t res, size_t i0,
                                size_t i1, size_t j0, size_t j1, real eps) {
    /* Each cell reads only its own velocity: updating in place is the
     * reference's collect-then-write. */
    for (size_t j = j0; j <= j1; j++)
        for (size_t i = i0; i <= i1; i++) {
            size_t idx = j * res + i;
            if (solid[idx] || h[idx] <= eps) continue;
            real ux = u[idx], vx = v[idx];
            for (int k = 0; k < 4; k++) {
                long ni = (long)i + NB4[k][0], nj = (long)j + NB4[k][1];
                if (!in_grid(ni, nj, res)) continue;
                if (!solid[(size_t)nj * res + (size_t)ni]) continue;
                real nx = (real)NB4[k][0], nz = (real)NB4[k][1];
                real vn = ux * nx + vx * nz;
                if (vn > K(0.0)) {
                    ux -= vn * nx;
                    vx -= vn * nz;
                }
            }
            u[idx] = ux;
            v[idx] = vx;
        }
}

/* displace_occupied, :4382-4444: water in cells the body occupies is split
 * over the open 4-neighbours (over the open cells within two cells when
 * none is open); a column with nowhere to go stays. */
void F(oracle_displace_occupied)(real *h, real *u, real *v, const uint8_t *solid, size_t res, real *give) {
    size_t n = res * res;
    for (size_t k = 0; k < n; k++) give[k] = K(0.0);
    for (size_t j = 0; j < res; j++)
        for (size_t i = 0; i < res; i++) {
            size_t idx = j * res + i;
            if (!solid[idx]) continue;
            real vol = h[idx];
            if (vol <= K(1.0e-8)) {
                h[idx] = u[idx] = v[idx] = K(0.0);
                continue;
            }
            size_t neigh[24];
            int nn = 0;
            for (int k = 0; k < 4; k++) {
                long ni = (long)i + NB4[k][0], nj = (long)j + NB4[k][1];
                if (!in_grid(ni, nj, res)) continue;
                size_t nidx = (size_t)nj * res + (size_t)ni;
                if (!solid[nidx]) neigh[nn++] = nidx;
            }
            if (nn == 0)
                for (int dj = -2; dj <= 2; dj++)
                    for (int di = -2; di <= 2; di++) {
                        long ni = (long)i + di, nj = (long)j + dj;
                        if (!in_grid(ni, nj, res)) continue;
                        size_t nidx = (size_t)nj * res + (size_t)ni;
                        if (!solid[nidx]) neigh[nn++] = nidx;
                    }
            if (nn == 0) continue;
            h[idx] = u[idx] = v[idx] = K(0.0);
            real share = vol / (real)nn;
            for (int q = 0; q < nn; q++) give[neigh[q]] += share;
        }
    for (size_t k = 0; k < n; k++)
        if (give[k] > K(0.0)) h[k] += give[k];
}

/* redirect_into_walls, :4446-4563: discharge aimed at a wall moves to the
 * open flanks beside it (books over the box and a 1-cell rim). */
void F(oracle_redirect_into_walls)(real *h, const real *u, const real *v, const uint8_t *solid, const real *bed,
                                   size_t res, real cell, real dt, size_t i0, size_t i1, size_t j0, size_t j1,
                                   real *removed, real *added) {
    size_t ri0 = i0 ? i0 - 1 : 0, ri1 = i1 + 1 < res - 1 ? i1 + 1 : res - 1;
    size_t rj0 = j0 ? j0 - 1 : 0, rj1 = j1 + 1 < res - 1 ? j1 + 1 : res - 1;
    for (size_t j = rj0; j <= rj1; j++)
        for (size_t i = ri0; i <= ri1; i++) removed[j * res + i] = added[j * res + i] = K(0.0);
    for (size_t j = j0; j <= j1; j++)
        for (size_t i = i0; i <= i1; i++) {
            size_t idx = j * res + i;
            if (solid[idx] || h[idx] < K(0.02)) continue;
            for (int k = 0; k < 4; k++) {
                int di = NB4[k][0], dj = NB4[k][1];
                long ni = (long)i + di, nj = (long)j + dj;
                if (!in_grid(ni, nj, res)) continue;
                if (!solid[(size_t)nj * res + (size_t)ni]) continue;
                real vn = u[idx] * (real)di + v[idx] * (real)dj;
                long ui = (long)i - di, uj = (long)j - dj;
                if (in_grid(ui, uj, res)) {
                    size_t uidx = (size_t)uj * res + (size_t)ui;
                    if (!solid[uidx]) {
                        real feed = u[uidx] * (real)di + v[uidx] * (real)dj;
                        if (feed > vn) vn = feed;
                    }
                }
                if (vn < K(0.05)) continue;
                int perps[2][2];
                if (di != 0) {
                    perps[0][0] = 0; perps[0][1] = -1; perps[1][0] = 0; perps[1][1] = 1;
                } else {
                    perps[0][0] = -1; perps[0][1] = 0; perps[1][0] = 1; perps[1][1] = 0;
                }
                size_t dest[2];
                int nd = 0;
                real flank = K(0.0);
                for (int p = 0; p < 2; p++) {
                    long qi = (long)i + perps[p][0], qj = (long)j + perps[p][1];
                    if (!in_grid(qi, qj, res)) continue;
                    size_t q = (size_t)qj * res + (size_t)qi;
                    if (solid[q]) continue;
                    if (h[q] <= K(1.0e-6) && bed[q] > bed[idx] + h[idx] + K(1.0e-3)) continue;
                    dest[nd++] = q;
                    flank += h[q];
                }
                if (nd == 0) continue;
                real excess = h[idx] - flank / (real)nd;
                if (excess < K(2.0e-3)) continue;
                real room = h[idx] - removed[idx];
                if (room < K(1.0e-6)) continue;
                real frac = R_FMIN(vn * dt / cell, K(0.45));
                real dh = R_FMIN(R_FMIN(excess, h[idx] * frac), room);
                if (dh < K(1.0e-8)) continue;
                removed[idx] += dh;
                real share = dh / (real)nd;
                for (int q = 0; q < nd; q++) added[dest[q]] += share;
            }
        }
    for (size_t j = rj0; j <= rj1; j++)
        for (size_t i = ri0; i <= ri1; i++) {
            size_t idx = j * res + i;
            if (removed[idx] > K(0.0) || added[idx] > K(0.0))
                h[idx] = R_FMAX(h[idx] - removed[idx] + added[idx], K(0.0));
        }
}

/* move_edge_bufs, :4178-4223 (solid and bed always given on this path). */
static inline void move_edge_walled(const real *h, const uint8_t *gate, size_t a, size_t b, real vel, real cell,
                                    real dt, const uint8_t *solid, const real *bed, real *removed, real *added) {
    if (solid[a] || solid[b]) return;
    if (R_FABS(vel) < K(1.0e-5)) return;
    size_t donor = vel > K(0.0) ? a : b, recv = vel > K(0.0) ? b : a;
    if (h[recv] <= K(1.0e-6) && bed[recv] > bed[donor] + h[donor] + K(1.0e-3)) return;
    real room = h[donor] - removed[donor];
    if (room <= K(1.0e-8)) return;
    if (gate[recv] == 0 && h[recv] <= K(1.0e-6)) return;
    real frac = R_FMIN(R_FABS(vel) * dt / cell, K(0.45));
    real dh = R_FMIN(h[donor] * frac, room);
    if (dh <= K(1.0e-8)) return;
    removed[donor] += dh;
    added[recv] += dh;
}

/* transport with solids, :3894-3977 (transport_par :3979-4148 walks the
 * same edges in the same order per row or column). */
void F(oracle_transport_walled)(real *h, const real *u, const real *v, const uint8_t *gate, const uint8_t *solid,
                                const real *bed, size_t res, real cell, real dt, size_t i0, size_t i1, size_t j0,
                                size_t j1, int along_x, real *removed, real *added) {
    for (size_t j = j0; j <= j1; j++)
        for (size_t i = i0; i <= i1; i++) removed[j * res + i] = added[j * res + i] = K(0.0);
    if (along_x) {
        for (size_t j = j0; j <= j1; j++)
            for (size_t i = i0; i < i1; i++) {
                size_t a = j * res + i;
                move_edge_walled(h, gate, a, a + 1, K(0.5) * (u[a] + u[a + 1]), cell, dt, solid, bed, removed,
                                 added);
            }
    } else {
        for (size_t j = j0; j < j1; j++)
            for (size_t i = i0; i <= i1; i++) {
                size_t a = j * res + i;
                move_edge_walled(h, gate, a, a + res, K(0.5) * (v[a] + v[a + res]), cell, dt, solid, bed,
                                 removed, added);
            }
    }
    Box b = {i0, i1, j0, j1};
    apply_moves(h, removed, added, res, b);
}

/*
 * Gravity beside walls, force_gravity_solid :2101-2128 (one pass: kick and
 * confinement cell by cell in raster order) and force_gravity_solid_par
 * :2131-2200 (every kick first, then the confinement sweep): the large-box
 * form with row splits allowed, as the reference picks it.  Wind on a walled
 * sheet: blend and confinement cell by cell (:1950-1964).
 */
static void force_walled(const real *h, real *u, real *v, const real *nu, const real *p, const uint8_t *mask,
                         size_t res, real cell, real dt, Drive drive, real eps, Box b, int inner_par) {
    if (drive.wind) {
        real blend = K(1.0) - R_EXP(-dt * K(1.6));
        for (size_t j = b.j0; j <= b.j1; j++)
            for (size_t i = b.i0; i <= b.i1; i++) {
                size_t idx = j * res + i;
                if (h[idx] <= eps) continue;
                u[idx] += (drive.wx - u[idx]) * blend;
                v[idx] += (drive.wz - v[idx]) * blend;
                if (drive.vorticity > K(0.0)) confine(u, v, u, v, res, i, j, cell, dt, drive.vorticity);
            }
        return;
    }
    int par = (b.j1 - b.j0 >= 47) && (b.i1 - b.i0 >= 31) && inner_par;
    for (size_t j = b.j0; j <= b.j1; j++)
        for (size_t i = b.i0; i <= b.i1; i++) {
            size_t idx = j * res + i;
            if (h[idx] <= eps) continue;
            real px, pz;
            grad(p, res, i, j, cell, &px, &pz);
            real mob = F(oracle_sheet_mobility)(nu[idx]);
            if (h[idx] > K(0.08) && solid_near(mask, res, i, j, 1)) mob = R_FMAX(mob, K(1.0));
            u[idx] += -GRAVITY * px * mob * dt;
            v[idx] += -GRAVITY * pz * mob * dt;
            if (!par && drive.vorticity > K(0.0)) confine(u, v, u, v, res, i, j, cell, dt, drive.vorticity);
        }
    if (!par || drive.vorticity <= K(0.0)) return;
    for (size_t j = b.j0; j <= b.j1; j++)
        for (size_t i = b.i0; i <= b.i1; i++) {
            if (h[j * res + i] <= eps) continue;
            confine(u, v, u, v, res, i, j, cell, dt, drive.vorticity);
        }
}

/* The substep with walls (:1879-2082, solid = Some).  Scratch e doubles as
 * the lift's and the ghosts' temporary, su / sv as the ghost velocities. */
static void substep_walled(real *h, real *u, real *v, const real *bed, const real *nu, const uint8_t *gate,
                           const uint8_t *mask, size_t res, real cell, real dt, Drive drive, real eps, Box b,
                           int inner_par, Scratch *sc) {
    size_t hi0 = b.i0 >= 2 ? b.i0 - 2 : 0, hi1 = (b.i1 + 2 < res - 1) ? b.i1 + 2 : res - 1;
    size_t hj0 = b.j0 >= 2 ? b.j0 - 2 : 0, hj1 = (b.j1 + 2 < res - 1) ? b.j1 + 2 : res - 1;
    for (size_t j = hj0; j <= hj1; j++)
        for (size_t i = hi0; i <= hi1; i++) {
            size_t k = j * res + i;
            sc->b[k] = sc->c[k] = sc->d[k] = sc->e[k] = K(0.0);
            sc->a[k] = bed[k] + h[k];
        }
    F(oracle_lift_solid_surface)(sc->a, h, mask, res, b.i0, b.i1, b.j0, b.j1, eps, sc->e);
    blur_box(sc->a, sc->b, sc->e, res, b, GAUSS_F);
    blur_box(sc->a, sc->c, sc->e, res, b, STEEP_F);
    for (size_t j = b.j0; j <= b.j1; j++)
        for (size_t i = b.i0; i <= b.i1; i++) {
            size_t idx = j * res + i;
            real filtered = K(0.62) * sc->b[idx] + K(0.38) * sc->c[idx];
            sc->d[idx] = K(0.45) * filtered + K(0.55) * sc->a[idx];
        }
    F(oracle_noslip_ghost)(u, v, h, mask, res, b.i0, b.i1, b.j0, b.j1, eps, sc->su, sc->sv);
    force_walled(h, u, v, nu, sc->d, mask, res, cell, dt, drive, eps, b, inner_par);
    /* The projection always runs beside walls (:1978-2002). */
    for (size_t j = b.j0; j <= b.j1; j++)
        for (size_t i = b.i0; i <= b.i1; i++) {
            size_t idx = j * res + i;
            size_t ip = i + 1 < res ? i + 1 : res - 1, im = i ? i - 1 : 0;
            size_t jp = j + 1 < res ? j + 1 : res - 1, jm = j ? j - 1 : 0;
            real ux = (u[j * res + ip] - u[j * res + im]) / (K(2.0) * cell);
            real vz = (v[jp * res + i] - v[jm * res + i]) / (K(2.0) * cell);
            real div = ux + vz;
            real lap_h = lap(sc->a, res, i, j, cell);
            real overload = K(1.0) + clampr(nu[idx] / K(231.0), K(0.0), K(6.0));
            sc->b[idx] = div * overload - K(0.015) * overload * lap_h;
        }
    blur_box(sc->b, sc->c, sc->e, res, b, GAUSS_F);
    blur_box(sc->b, sc->a, sc->e, res, b, STEEP_F);
    for (size_t j = b.j0; j <= b.j1; j++)
        for (size_t i = b.i0; i <= b.i1; i++) {
            size_t idx = j * res + i;
            sc->d[idx] = K(0.62) * sc->c[idx] + K(0.38) * sc->a[idx];
        }
    real lim = drive.wind ? R_FMAX(R_HYPOT(drive.wx, drive.wz), FILM_SPEED_CAP) : FILM_SPEED_CAP;
    for (size_t j = b.j0; j <= b.j1; j++) {
        int memo_set = 0;
        real memo_nu = K(0.0), memo_decay = K(1.0);
        for (size_t i = b.i0; i <= b.i1; i++) {
            size_t idx = j * res + i;
            if (h[idx] <= eps) continue;
            real qx, qz;
            grad(sc->d, res, i, j, cell, &qx, &qz);
            u[idx] += K(0.012) * dt * qx;
            v[idx] += K(0.012) * dt * qz;
            real nu_here = nu[idx];
            if (!memo_set || memcmp(&nu_here, &memo_nu, sizeof(real)) != 0) {
                memo_set = 1;
                memo_nu = nu_here;
                memo_decay = R_EXP(-F(oracle_sheet_damp)(nu_here) * dt);
            }
            u[idx] = finite_speed(u[idx] * memo_decay, lim);
            v[idx] = finite_speed(v[idx] * memo_decay, lim);
        }
    }
    F(oracle_redirect_into_walls)(h, u, v, mask, bed, res, cell, dt, b.i0, b.i1, b.j0, b.j1, sc->removed,
                                  sc->added);
    F(oracle_block_into_solid)(u, v, h, mask, res, b.i0, b.i1, b.j0, b.j1, eps);
    F(oracle_noslip_ghost)(u, v, h, mask, res, b.i0, b.i1, b.j0, b.j1, eps, sc->su, sc->sv);
    F(oracle_transport_walled)(h, u, v, gate, mask, bed, res, cell, dt, b.i0, b.i1, b.j0, b.j1, 1, sc->removed,
                               sc->added);
    F(oracle_transport_walled)(h, u, v, gate, mask, bed, res, cell, dt, b.i0, b.i1, b.j0, b.j1, 0, sc->removed,
                               sc->added);
    for (size_t j = b.j0; j <= b.j1; j++)
        for (size_t i = b.i0; i <= b.i1; i++) {
            size_t idx = j * res + i;
            if (eps > K(0.0) && h[idx] < eps) h[idx] = u[idx] = v[idx] = K(0.0);
        }
}

/*
 * The step: step_sheet_eps (:856-889) and step_sheet_scratch (:1659-1831).
 * Returns the number of substeps taken (0 when the call is a no-op).
 * boxes_out (optional, 8 x 4 entries) receives the work boxes.
 */
int F(oracle_step_sheet_ex)(real *h, real *u, real *v, const real *bed, const real *nu,
                            const uint8_t *gate, size_t res, real cell, real dt, int drive_wind,
                            real wind_x, real wind_z, real vorticity, real eps, int banks,
                            real sub_cap, int pace, int inner_par, int confine_mode,
                            size_t *boxes_out, size_t *nboxes_out) {
    if (nboxes_out) *nboxes_out = 0;
    if (!(dt > K(0.0)) || res == 0) return 0;
    size_t n = res * res;
    real sum = wet_sum(h, n, eps);
    if (sum <= eps) return 0;
    Drive drive = {drive_wind, wind_x, wind_z, vorticity};
    int open_fine = !drive_wind && cell <= FINE_CELL_M + K(1.0e-6);
    int fine_open = open_fine && pace == 0;
    int whole_wet = open_fine && (pace == 2 || pace == 3);
    int awake = open_fine && !whole_wet && !fine_open;
    size_t raw[4 * 4096];
    size_t nb;
    if (fine_open || awake)
        nb = F(oracle_awake_work_boxes)(h, u, v, bed, res, cell, eps, raw, 4096);
    else
        nb = F(oracle_sheet_work_boxes)(h, res, eps, 1, raw, 4096);
    if (nb > 4096) nb = 4096;
    if (nboxes_out) *nboxes_out = nb;
    if (boxes_out)
        for (size_t k = 0; k < nb && k < 8; k++)
            for (int q = 0; q < 4; q++) boxes_out[4 * k + q] = raw[4 * k + q];
    if (nb == 0) return 0;
    real cap = R_ISFINITE(sub_cap) ? clampr(sub_cap, K(1.0), K(8.0)) : K(8.0);
    uint32_t nsub;
    if (fine_open) {
        nsub = 1;
    } else {
        real max_s = K(0.15);
        if (drive_wind) max_s = R_FMAX(max_s, R_HYPOT(wind_x, wind_z));
        real max_s2 = K(0.0);
        for (size_t k = 0; k < nb; k++)
            for (size_t j = raw[4 * k + 2]; j <= raw[4 * k + 3]; j++)
                for (size_t i = raw[4 * k + 0]; i <= raw[4 * k + 1]; i++) {
                    real uu = u[j * res + i], vv = v[j * res + i];
                    real s2 = uu * uu + vv * vv;
                    if (s2 > max_s2) max_s2 = s2;
                }
        max_s = R_FMAX(max_s, R_SQRT(max_s2));
        real courant = drive_wind ? K(0.85) : K(0.42);
        real q = R_CEIL((dt * max_s) / (courant * R_FMAX(cell, K(1.0e-3))));
        nsub = (uint32_t)clampr(q, K(1.0), cap);
    }
    real dts = dt / (real)nsub;
    /* Nine scratch planes, kept per thread between calls as the reference
     * keeps its thread-local SCRATCH (:581-589): every plane a substep reads
     * is cleared or written first on the cells it reads. */
    static _Thread_local real *pool = NULL;
    static _Thread_local size_t pool_n = 0;
    if (pool_n < n) {
        free(pool);
        pool = calloc(9 * n, sizeof(real));
        pool_n = n;
    }
    Scratch sc;
    sc.a = pool; sc.b = pool + n; sc.c = pool + 2 * n; sc.d = pool + 3 * n; sc.e = pool + 4 * n;
    sc.removed = pool + 5 * n; sc.added = pool + 6 * n; sc.su = pool + 7 * n; sc.sv = pool + 8 * n;
    if (pace == 3 && whole_wet && nsub > 1) {
        for (size_t k = 0; k < nb; k++) {
            Box b = {raw[4 * k], raw[4 * k + 1], raw[4 * k + 2], raw[4 * k + 3]};
            substep(h, u, v, bed, nu, gate, res, cell, dt, drive, eps, b, banks, 0, inner_par,
                    confine_mode, &sc);
            pace3_slices(h, u, v, gate, res, cell, dts, nsub, eps, b, &sc);
        }
    } else {
        for (uint32_t s = 0; s < nsub; s++)
            for (size_t k = 0; k < nb; k++) {
                Box b = {raw[4 * k], raw[4 * k + 1], raw[4 * k + 2], raw[4 * k + 3]};
                substep(h, u, v, bed, nu, gate, res, cell, dts, drive, eps, b, banks, 1, inner_par,
                        confine_mode, &sc);
            }
    }
    return (int)nsub;
}


/*
 * The walled step: step_sheet_eps_in (:856-889) with solid = Some and the
 * plan of step_sheet_scratch (:1659-1831) for it: wet-set work boxes, every
 * substep at Courant 0.42 (wind too), no pace policy.  solid: res * res
 * bytes, nonzero = wall.  Returns the substep count.
 */
int F(oracle_step_sheet_walled)(real *h, real *u, real *v, const real *bed, const real *nu, const uint8_t *gate,
                                const uint8_t *solid, size_t res, real cell, real dt, int drive_wind, real wind_x,
                                real wind_z, real vorticity, real eps, real sub_cap, int inner_par) {
    if (!(dt > K(0.0)) || res == 0) return 0;
    size_t n = res * res;
    real sum = wet_sum(h, n, eps);
    if (sum <= eps) return 0;
    Drive drive = {drive_wind, wind_x, wind_z, vorticity};
    size_t raw[4 * 4096];
    size_t nb = F(oracle_sheet_work_boxes)(h, res, eps, 1, raw, 4096);
    if (nb > 4096) nb = 4096;
    if (nb == 0) return 0;
    real cap = R_ISFINITE(sub_cap) ? clampr(sub_cap, K(1.0), K(8.0)) : K(8.0);
    real max_s = K(0.15);
    if (drive_wind) max_s = R_FMAX(max_s, R_HYPOT(wind_x, wind_z));
    real max_s2 = K(0.0);
    for (size_t k = 0; k < nb; k++)
        for (size_t j = raw[4 * k + 2]; j <= raw[4 * k + 3]; j++)
            for (size_t i = raw[4 * k + 0]; i <= raw[4 * k + 1]; i++) {
                real uu = u[j * res + i], vv = v[j * res + i];
                real s2 = uu * uu + vv * vv;
                if (s2 > max_s2) max_s2 = s2;
            }
    max_s = R_FMAX(max_s, R_SQRT(max_s2));
    real q = R_CEIL((dt * max_s) / (K(0.42) * R_FMAX(cell, K(1.0e-3))));
    uint32_t nsub = (uint32_t)clampr(q, K(1.0), cap);
    real dts = dt / (real)nsub;
    real *pool = calloc(9 * n, sizeof(real));
    if (!pool) return -1;
    Scratch sc;
    sc.a = pool; sc.b = pool + n; sc.c = pool + 2 * n; sc.d = pool + 3 * n; sc.e = pool + 4 * n;
    sc.removed = pool + 5 * n; sc.added = pool + 6 * n; sc.su = pool + 7 * n; sc.sv = pool + 8 * n;
    for (uint32_t s = 0; s < nsub; s++)
        for (size_t k = 0; k < nb; k++) {
            Box b = {raw[4 * k], raw[4 * k + 1], raw[4 * k + 2], raw[4 * k + 3]};
            substep_walled(h, u, v, bed, nu, gate, solid, res, cell, dts, drive, eps, b, inner_par, &sc);
        }
    free(pool);
    return (int)nsub;
}

#ifdef ORACLE_THREADS
int F(oracle_step_sheet)(real *h, real *u, real *v, const real *bed, const real *nu, const uint8_t *gate,
                         size_t res, real cell, real dt, int drive_wind, real wind_x, real wind_z, real vorticity);

/*
 * The CPU baseline of one two-layer step: water and slurry as two tasks (the
 * reference's rayon::join in step_pair, :8127-8131), each a step_sheet
 * (:770-785) whose passes run in row stripes.  Returns the larger substep
 * count.
 */
int F(oracle_step_pair_par)(real *hw, real *uw, real *vw, const real *nu_w, real vort_w, real *hs, real *us,
                            real *vs, const real *nu_s, real vort_s, const real *bed, const uint8_t *gate, size_t res,
                            real cell, real dt) {
    int n0 = 0, n1 = 0;
    omp_set_max_active_levels(2);
#pragma omp parallel sections num_threads(2)
    {
#pragma omp section
        n0 = F(oracle_step_sheet)(hw, uw, vw, bed, nu_w, gate, res, cell, dt, 0, K(0.0), K(0.0), vort_w);
#pragma omp section
        n1 = F(oracle_step_sheet)(hs, us, vs, bed, nu_s, gate, res, cell, dt, 0, K(0.0), K(0.0), vort_s);
    }
    return n0 > n1 ? n0 : n1;
}
#endif

/*
 * Test entry points: one part of the substep on its own, through the same
 * functions the step calls.  oracle_surface_pressure writes the pressure
 * plane d of box (i0..i1, j0..j1) into d_out (box cells only);
 * oracle_force_open runs the gravity kick from the given pressure plane and
 * the confinement sweep on the wet cells of the box (confine_mode as in
 * oracle_step_sheet_ex, inner_par: row splits allowed).
 */
static real *scratch_pool(Scratch *sc, size_t n) {
    real *pool = calloc(9 * n, sizeof(real));
    sc->a = pool; sc->b = pool + n; sc->c = pool + 2 * n; sc->d = pool + 3 * n; sc->e = pool + 4 * n;
    sc->removed = pool + 5 * n; sc->added = pool + 6 * n; sc->su = pool + 7 * n; sc->sv = pool + 8 * n;
    return pool;
}

void F(oracle_surface_pressure)(const real *h, const real *bed, size_t res, size_t i0, size_t i1, size_t j0,
                                size_t j1, real *d_out) {
    Scratch sc;
    real *pool = scratch_pool(&sc, res * res);
    Box b = {i0, i1, j0, j1};
    surface_pressure(h, bed, res, b, 0, &sc);
    for (size_t j = j0; j <= j1; j++)
        for (size_t i = i0; i <= i1; i++) d_out[j * res + i] = sc.d[j * res + i];
    free(pool);
}

void F(oracle_force_open)(const real *h, real *u, real *v, const real *nu, const real *p, size_t res, real cell,
                          real dt, real vorticity, real eps, size_t i0, size_t i1, size_t j0, size_t j1,
                          int inner_par, int confine_mode) {
    Scratch sc;
    real *pool = scratch_pool(&sc, res * res);
    Box b = {i0, i1, j0, j1};
    Drive drive = {0, K(0.0), K(0.0), vorticity};
    force_open(h, u, v, nu, p, res, cell, dt, drive, eps, b, inner_par, confine_mode, &sc);
    free(pool);
}

/* step_sheet, src/fluid_core.rs:770-785: eps 1e-6, no banks, cap 8, pace 0. */
int F(oracle_step_sheet)(real *h, real *u, real *v, const real *bed, const real *nu,
                         const uint8_t *gate, size_t res, real cell, real dt, int drive_wind,
                         real wind_x, real wind_z, real vorticity) {
    return F(oracle_step_sheet_ex)(h, u, v, bed, nu, gate, res, cell, dt, drive_wind, wind_x, wind_z,
                                   vorticity, K(1.0e-6), 0, K(8.0), 0, 1, CONFINE_REFERENCE, NULL,
                                   NULL);
}

/* column_sum, src/fluid_core.rs:4617-4620 (sequential sum of max(h,0)). */
real F(oracle_column_sum)(const real *h, size_t n, real cell) {
    real area = cell * cell;
    real s = -K(0.0);
    for (size_t i = 0; i < n; i++) s += R_FMAX(h[i], K(0.0));
    return s * area;
}
