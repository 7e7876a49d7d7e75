/*
 * oracle.c -- CPU restatement of vibeSpatial's spatial-join path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
 * entry and the CPU-baseline leg of bench.py may load liboracle.so.  The product
 * (vibespatial_amd/) never links, imports or calls it.
 *
 * What it restates (citations are reference paths, file:line):
 *   - point location 0/1/2 with the half-open crossing rule and the
 *     boundary-in-closed-bbox rule:
 *       src/vibespatial/cuda/device_functions/point_in_ring.py:148-188
 *   - ring XOR over shell+holes, first part that yields boundary/interior wins:
 *       src/vibespatial/predicates/point_relations_kernels.py:37-95
 *   - orientation sign convention (a, b, c) -> (ax-cx)(by-cy) - (ay-cy)(bx-cx):
 *       src/vibespatial/cuda/device_functions/orient2d.py:37-84
 *     but EXACT: the reference adds the rounded tail B3 and is not a full
 *     expansion; here every tail of every difference and product is kept
 *     (Dekker/Knuth error-free transforms, Shewchuk 1997 expansion sums).
 *   - segment touch/cross with collinear containment:
 *       src/vibespatial/predicates/polygon_kernels.py:131-161
 *   - polygon intersects = any segment pair touches, else any vertex of A is
 *     located in B (code != 0), else any vertex of B in A:
 *       src/vibespatial/predicates/polygon_kernels.py:163-228
 *   - inclusive bbox overlap, NaN rows never match:
 *       src/vibespatial/kernels/core/spatial_query_source.py:441-555
 *   - 16-level (32-bit) Hilbert encoder, public-domain prefix-scan method:
 *       src/vibespatial/api/tools/hilbert_curve.py:189-247
 *   - predicate mapping of location codes:
 *       src/vibespatial/predicates/point_relations.py:100-129
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   or_orient2d_exact   vs Python fractions.Fraction on random and adversarial
 *                       near-collinear triples (exact rationals are ground truth)
 *   or_orient2d         (filtered) vs or_orient2d_exact on the same triples
 *   or_point_location   reference literal vectors (tests/golden/), Fraction
 *                       restatement on tiny polygons, invariants (vertex and
 *                       dyadic-midpoint points are boundary, centre of a star
 *                       polygon is interior, translation by powers of two)
 *   or_segments_touch   Fraction brute force; textbook cases
 *   or_polygon_intersects  golden cases; brute force by symmetric definition
 *   or_hilbert16        textbook rotate/flip xy2d on full 2^k grids (bijective,
 *                       4-neighbour adjacency of consecutive codes)
 *   or_bbox_pairs       numpy broadcast brute force
 *   or_sjoin_points_polygons / or_sjoin_polygons_polygons
 *                       == bbox brute force + location on every pair
 *
 * Coordinates are fp64 SoA; polygonal rows use GeoArrow-style offsets:
 *   geom_offsets[row..row+1]  -> parts, part_offsets[part..part+1] -> rings,
 *   ring_offsets[ring..ring+1] -> coordinates (rings closed, first == last).
 * Exactness assumes no overflow/underflow in products of coordinate
 * differences (|coords| well inside 2^+-500), as Shewchuk's predicates do.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  Contraction
 * must stay off: the error-free transforms rely on separately rounded ops.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Error-free transforms and expansion sums                            */
/* ------------------------------------------------------------------ */

static inline void ef_two_sum(double a, double b, double *s, double *e) {
    double x = a + b;
    double bv = x - a;
    double av = x - bv;
    double br = b - bv;
    double ar = a - av;
    *s = x;
    *e = ar + br;
}

static inline void ef_two_diff(double a, double b, double *d, double *e) {
    double x = a - b;
    double bv = a - x;
    double av = x + bv;
    double br = bv - b;
    double ar = a - av;
    *d = x;
    *e = ar + br;
}

static inline void ef_two_prod(double a, double b, double *p, double *e) {
    double x = a * b;
    *p = x;
    *e = fma(a, b, -x); /* exact product tail (correctly rounded fma) */
}

/* Add b to the nonoverlapping, magnitude-increasing expansion e[0..n) in place
 * (Shewchuk's GROW-EXPANSION with zero elimination). Returns the new length. */
static int ef_grow(double *e, int n, double b) {
    double q = b;
    int m = 0;
    for (int i = 0; i < n; ++i) {
        double s, h;
        ef_two_sum(q, e[i], &s, &h);
        q = s;
        if (h != 0.0) e[m++] = h;
    }
    if (q != 0.0 || m == 0) e[m++] = q;
    return m;
}

/* Exact sign of (ax-cx)(by-cy) - (ay-cy)(bx-cx); no filter. */
int or_orient2d_exact(double ax, double ay, double bx, double by, double cx, double cy) {
    double acx, acxt, bcx, bcxt, acy, acyt, bcy, bcyt;
    ef_two_diff(ax, cx, &acx, &acxt);
    ef_two_diff(bx, cx, &bcx, &bcxt);
    ef_two_diff(ay, cy, &acy, &acyt);
    ef_two_diff(by, cy, &bcy, &bcyt);
    const double l[4][2] = {{acx, bcy}, {acx, bcyt}, {acxt, bcy}, {acxt, bcyt}};
    const double r[4][2] = {{acy, bcx}, {acy, bcxt}, {acyt, bcx}, {acyt, bcxt}};
    double e[40];
    int n = 0;
    for (int k = 0; k < 4; ++k) {
        double p, t;
        ef_two_prod(l[k][0], l[k][1], &p, &t);
        n = ef_grow(e, n, t);
        n = ef_grow(e, n, p);
        ef_two_prod(r[k][0], r[k][1], &p, &t);
        n = ef_grow(e, n, -t);
        n = ef_grow(e, n, -p);
    }
    const double top = e[n - 1];
    return (top > 0.0) - (top < 0.0);
}

/* Filtered: Shewchuk's stage-A bound (3 + 16 eps) eps, eps = 2^-53, then exact. */
int or_orient2d(double ax, double ay, double bx, double by, double cx, double cy) {
    const double dl = (ax - cx) * (by - cy);
    const double dr = (ay - cy) * (bx - cx);
    const double det = dl - dr;
    const double bound = 3.3306690738754716e-16 * (fabs(dl) + fabs(dr));
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return or_orient2d_exact(ax, ay, bx, by, cx, cy);
}

/* ------------------------------------------------------------------ */
/* Point location                                                      */
/* ------------------------------------------------------------------ */

/* 0 outside, 1 boundary, 2 inside, for one ring x[start..end). */
static int ring_classify(double px, double py, const double *x, const double *y,
                         int64_t start, int64_t end) {
    if (end - start < 2) return 0;
    int inside = 0;
    for (int64_t i = start, j = end - 1; i < end; j = i++) {
        const double ax = x[j], ay = y[j], bx = x[i], by = y[i];
        const int crosses = (ay > py) != (by > py);
        const double minx = fmin(ax, bx), maxx = fmax(ax, bx);
        const int in_box = py >= fmin(ay, by) && py <= fmax(ay, by) && px >= minx && px <= maxx;
        if (!crosses && !in_box) continue;
        if (crosses && px < minx) { inside ^= 1; continue; }
        if (crosses && px > maxx) continue;
        const int o = or_orient2d(ax, ay, bx, by, px, py);
        if (in_box && o == 0) return 1;
        if (crosses && ((o > 0) == (by > ay))) inside ^= 1;
    }
    return inside ? 2 : 0;
}

int or_point_location(double px, double py, const double *x, const double *y,
                      const int32_t *geom_offsets, const int32_t *part_offsets,
                      const int32_t *ring_offsets, int64_t row) {
    if (isnan(px) || isnan(py)) return 0;
    for (int32_t part = geom_offsets[row]; part < geom_offsets[row + 1]; ++part) {
        int inside = 0;
        for (int32_t ring = part_offsets[part]; ring < part_offsets[part + 1]; ++ring) {
            const int k = ring_classify(px, py, x, y, ring_offsets[ring], ring_offsets[ring + 1]);
            if (k == 1) return 1;
            if (k == 2) inside ^= 1;
        }
        if (inside) return 2;
    }
    return 0;
}

/* Batch: codes[i] = location of (px[li[i]], py[li[i]]) in polygon row ri[i]. */
void or_point_location_pairs(const double *px, const double *py, const int64_t *li,
                             const int64_t *ri, int64_t n, const double *x, const double *y,
                             const int32_t *geom_offsets, const int32_t *part_offsets,
                             const int32_t *ring_offsets, uint8_t *codes, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1024)
#endif
    for (int64_t i = 0; i < n; ++i)
        codes[i] = (uint8_t)or_point_location(px[li[i]], py[li[i]], x, y, geom_offsets,
                                              part_offsets, ring_offsets, ri[i]);
}

/* ------------------------------------------------------------------ */
/* Segment and polygon intersects                                      */
/* ------------------------------------------------------------------ */

static inline int on_segment_collinear(double px, double py, double ax, double ay, double bx,
                                       double by) {
    return px >= fmin(ax, bx) && px <= fmax(ax, bx) && py >= fmin(ay, by) && py <= fmax(ay, by);
}

int or_segments_touch(double ax, double ay, double bx, double by, double cx, double cy,
                      double dx, double dy) {
    if (fmax(ax, bx) < fmin(cx, dx) || fmax(cx, dx) < fmin(ax, bx) ||
        fmax(ay, by) < fmin(cy, dy) || fmax(cy, dy) < fmin(ay, by))
        return 0;
    const int o1 = or_orient2d(ax, ay, bx, by, cx, cy);
    const int o2 = or_orient2d(ax, ay, bx, by, dx, dy);
    const int o3 = or_orient2d(cx, cy, dx, dy, ax, ay);
    const int o4 = or_orient2d(cx, cy, dx, dy, bx, by);
    if (o1 * o2 < 0 && o3 * o4 < 0) return 1;
    if (o1 == 0 && on_segment_collinear(cx, cy, ax, ay, bx, by)) return 1;
    if (o2 == 0 && on_segment_collinear(dx, dy, ax, ay, bx, by)) return 1;
    if (o3 == 0 && on_segment_collinear(ax, ay, cx, cy, dx, dy)) return 1;
    if (o4 == 0 && on_segment_collinear(bx, by, cx, cy, dx, dy)) return 1;
    return 0;
}

typedef struct {
    const double *x, *y;
    const int32_t *geom, *part, *ring;
} or_polys;

/* vertices tested for containment skip the closing duplicate (polygon_kernels.py:199-225) */
static inline int64_t last_vertex(int64_t c0, int64_t c1) { return c1 > c0 + 1 ? c1 - 1 : c1; }

static int poly_intersects(const or_polys *A, int64_t ra, const or_polys *B, int64_t rb) {
    const int32_t ap0 = A->geom[ra], ap1 = A->geom[ra + 1];
    const int32_t bp0 = B->geom[rb], bp1 = B->geom[rb + 1];
    const int64_t ac0 = A->ring[A->part[ap0]], ac1 = A->ring[A->part[ap1]];
    const int64_t bc0 = B->ring[B->part[bp0]], bc1 = B->ring[B->part[bp1]];
    if (ac1 <= ac0 || bc1 <= bc0) return 0; /* empty never intersects */
    /* 1. any segment pair touches or crosses */
    for (int32_t ar = A->part[ap0]; ar < A->part[ap1]; ++ar)
        for (int64_t ai = A->ring[ar] + 1; ai < A->ring[ar + 1]; ++ai)
            for (int32_t br = B->part[bp0]; br < B->part[bp1]; ++br)
                for (int64_t bi = B->ring[br] + 1; bi < B->ring[br + 1]; ++bi)
                    if (or_segments_touch(A->x[ai - 1], A->y[ai - 1], A->x[ai], A->y[ai],
                                          B->x[bi - 1], B->y[bi - 1], B->x[bi], B->y[bi]))
                        return 1;
    /* 2. a vertex of A located in B, 3. a vertex of B located in A */
    for (int32_t ar = A->part[ap0]; ar < A->part[ap1]; ++ar)
        for (int64_t v = A->ring[ar]; v < last_vertex(A->ring[ar], A->ring[ar + 1]); ++v)
            if (or_point_location(A->x[v], A->y[v], B->x, B->y, B->geom, B->part, B->ring, rb))
                return 1;
    for (int32_t br = B->part[bp0]; br < B->part[bp1]; ++br)
        for (int64_t v = B->ring[br]; v < last_vertex(B->ring[br], B->ring[br + 1]); ++v)
            if (or_point_location(B->x[v], B->y[v], A->x, A->y, A->geom, A->part, A->ring, ra))
                return 1;
    return 0;
}

int or_polygon_intersects(const double *ax, const double *ay, const int32_t *ag,
                          const int32_t *ap, const int32_t *ar, int64_t ra, const double *bx,
                          const double *by, const int32_t *bg, const int32_t *bp,
                          const int32_t *br, int64_t rb) {
    or_polys A = {ax, ay, ag, ap, ar}, B = {bx, by, bg, bp, br};
    return poly_intersects(&A, ra, &B, rb);
}

/* ------------------------------------------------------------------ */
/* Bounds and bbox pairs                                               */
/* ------------------------------------------------------------------ */

void or_polygon_bounds(const double *x, const double *y, const int32_t *geom,
                       const int32_t *part, const int32_t *ring, int64_t n, double *out) {
    for (int64_t r = 0; r < n; ++r) {
        const int64_t c0 = ring[part[geom[r]]], c1 = ring[part[geom[r + 1]]];
        double b[4] = {NAN, NAN, NAN, NAN};
        if (c1 > c0) {
            b[0] = b[2] = x[c0];
            b[1] = b[3] = y[c0];
            for (int64_t c = c0 + 1; c < c1; ++c) {
                b[0] = fmin(b[0], x[c]); b[1] = fmin(b[1], y[c]);
                b[2] = fmax(b[2], x[c]); b[3] = fmax(b[3], y[c]);
            }
        }
        memcpy(out + 4 * r, b, sizeof b);
    }
}

static inline int box_overlap(const double *q, const double *t) {
    return q[0] <= t[2] && q[2] >= t[0] && q[1] <= t[3] && q[3] >= t[1];
}

/* Brute-force inclusive bbox pairs, lexsorted by (query, tree).  Pass out_* NULL
 * to count only; returns the pair count. */
int64_t or_bbox_pairs(const double *qb, int64_t nq, const double *tb, int64_t nt,
                      int64_t *out_q, int64_t *out_t) {
    int64_t k = 0;
    for (int64_t i = 0; i < nq; ++i)
        for (int64_t j = 0; j < nt; ++j)
            if (box_overlap(qb + 4 * i, tb + 4 * j)) {
                if (out_q) { out_q[k] = i; out_t[k] = j; }
                ++k;
            }
    return k;
}

/* ------------------------------------------------------------------ */
/* Hilbert encoders                                                    */
/* ------------------------------------------------------------------ */

static inline uint32_t spread16(uint32_t v) {
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

/* 16-level encoder over 16-bit integer coordinates; restates the prefix-scan
 * method of hilbert_curve.py:189-247 for level == 16. */
uint32_t or_hilbert16(uint32_t x, uint32_t y) {
    x &= 0xFFFFu; y &= 0xFFFFu;
    uint32_t a = x ^ y, b = 0xFFFFu ^ a, c = 0xFFFFu ^ (x | y), d = x & (y ^ 0xFFFFu);
    uint32_t A = a | (b >> 1), B = (a >> 1) ^ a;
    uint32_t C = ((c >> 1) ^ (b & (d >> 1))) ^ c, D = ((a & (c >> 1)) ^ (d >> 1)) ^ d;
    for (int s = 2; s <= 8; s <<= 1) {
        a = A; b = B; c = C; d = D;
        if (s < 8) {
            A = (a & (a >> s)) ^ (b & (b >> s));
            B = (a & (b >> s)) ^ (b & ((a ^ b) >> s));
        }
        C ^= (a & (c >> s)) ^ (b & (d >> s));
        D ^= (b & (c >> s)) ^ ((a ^ b) & (d >> s));
    }
    a = C ^ (C >> 1);
    b = D ^ (D >> 1);
    const uint32_t i0 = x ^ y;
    const uint32_t i1 = (b | (0xFFFFu ^ (i0 | a))) & 0xFFFFu;
    return (spread16(i1) << 1) | spread16(i0);
}

void or_hilbert16_batch(const uint32_t *x, const uint32_t *y, int64_t n, uint32_t *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = or_hilbert16(x[i], y[i]);
}

/* Textbook rotate/flip Hilbert index on a 2^order grid: an independent
 * algorithm used only to check or_hilbert16 (up to the curve's orientation). */
uint64_t or_hilbert_textbook(int order, uint64_t x, uint64_t y) {
    uint64_t d = 0;
    for (uint64_t s = (uint64_t)1 << (order - 1); s > 0; s >>= 1) {
        const uint64_t rx = (x & s) > 0, ry = (y & s) > 0;
        d += s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) { x = s - 1 - x; y = s - 1 - y; }
            const uint64_t t = x; x = y; y = t;
        }
    }
    return d;
}

/* ------------------------------------------------------------------ */
/* CPU join (the timed CPU baseline and the full-size parity checker)  */
/* ------------------------------------------------------------------ */
/* A packed tree over polygon boxes: rows sorted by the 16-level Hilbert key of
 * their box centre, leaves of `fan` boxes, parents built bottom up (same design
 * as io/gpu_parse/indexing.py:262-400).  Point queries descend it recursively;
 * the exact location is evaluated on each surviving pair.  OpenMP over
 * contiguous chunks of points; per-thread outputs are concatenated in thread
 * order so the result is lexsorted without a global sort. */

typedef struct {
    int64_t n_items, n_levels;
    int64_t level_start[24], level_count[24]; /* level 0 = root ... last = items */
    double *box;                              /* all levels, 4 doubles per node */
    int64_t *item_row;                        /* original row of each item slot */
    int fan;
} or_tree;

static int cmp_u64pair(const void *a, const void *b) {
    const uint64_t *x = (const uint64_t *)a, *y = (const uint64_t *)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    return (x[1] > y[1]) - (x[1] < y[1]);
}

static void tree_build(or_tree *t, const double *bounds, int64_t n, int fan) {
    memset(t, 0, sizeof *t);
    t->fan = fan;
    t->n_items = n;
    double ext[4] = {INFINITY, INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i) {
        const double *b = bounds + 4 * i;
        if (isnan(b[0])) continue;
        const double cx = 0.5 * (b[0] + b[2]), cy = 0.5 * (b[1] + b[3]);
        ext[0] = fmin(ext[0], cx); ext[1] = fmin(ext[1], cy);
        ext[2] = fmax(ext[2], cx); ext[3] = fmax(ext[3], cy);
    }
    uint64_t *kv = (uint64_t *)malloc(sizeof(uint64_t) * 2 * (n ? n : 1));
    const double sx = ext[2] > ext[0] ? 65535.0 / (ext[2] - ext[0]) : 0.0;
    const double sy = ext[3] > ext[1] ? 65535.0 / (ext[3] - ext[1]) : 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double *b = bounds + 4 * i;
        uint64_t key = UINT64_MAX;
        if (!isnan(b[0])) {
            const double cx = 0.5 * (b[0] + b[2]), cy = 0.5 * (b[1] + b[3]);
            key = or_hilbert16((uint32_t)((cx - ext[0]) * sx), (uint32_t)((cy - ext[1]) * sy));
        }
        kv[2 * i] = key;
        kv[2 * i + 1] = (uint64_t)i;
    }
    qsort(kv, (size_t)n, 2 * sizeof(uint64_t), cmp_u64pair);
    /* level sizes, items last */
    int64_t counts[24];
    int nl = 0;
    int64_t c = n;
    counts[nl++] = c;
    while (c > 1) { c = (c + fan - 1) / fan; counts[nl++] = c; }
    if (n == 0) { counts[0] = 0; nl = 1; }
    t->n_levels = nl;
    int64_t total = 0;
    for (int l = 0; l < nl; ++l) {
        t->level_count[l] = counts[nl - 1 - l];
        t->level_start[l] = total;
        total += t->level_count[l];
    }
    t->box = (double *)malloc(sizeof(double) * 4 * (total ? total : 1));
    t->item_row = (int64_t *)malloc(sizeof(int64_t) * (n ? n : 1));
    double *items = t->box + 4 * t->level_start[nl - 1];
    for (int64_t i = 0; i < n; ++i) {
        t->item_row[i] = (int64_t)kv[2 * i + 1];
        memcpy(items + 4 * i, bounds + 4 * kv[2 * i + 1], 4 * sizeof(double));
    }
    free(kv);
    for (int l = nl - 2; l >= 0; --l) {
        const double *child = t->box + 4 * t->level_start[l + 1];
        double *node = t->box + 4 * t->level_start[l];
        for (int64_t p = 0; p < t->level_count[l]; ++p) {
            double b[4] = {INFINITY, INFINITY, -INFINITY, -INFINITY};
            for (int64_t k = p * fan; k < (p + 1) * fan && k < t->level_count[l + 1]; ++k) {
                const double *cb = child + 4 * k;
                if (isnan(cb[0])) continue;
                b[0] = fmin(b[0], cb[0]); b[1] = fmin(b[1], cb[1]);
                b[2] = fmax(b[2], cb[2]); b[3] = fmax(b[3], cb[3]);
            }
            if (b[0] > b[2]) b[0] = b[1] = b[2] = b[3] = NAN;
            memcpy(node + 4 * p, b, sizeof b);
        }
    }
}

static void tree_free(or_tree *t) { free(t->box); free(t->item_row); }

typedef struct { int64_t *v; int64_t n, cap; } i64vec;
static void vpush(i64vec *v, int64_t a) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 1024;
        v->v = (int64_t *)realloc(v->v, sizeof(int64_t) * v->cap);
    }
    v->v[v->n++] = a;
}

static int cmp_i64(const void *a, const void *b) {
    const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* Collect the rows whose box contains box q (inclusive) into hits. */
static void tree_query(const or_tree *t, const double *q, i64vec *hits) {
    int64_t stack[24 * 64];
    int sl[24 * 64];
    int sp = 0;
    if (t->n_items == 0) return;
    stack[sp] = 0; sl[sp] = 0; ++sp;
    const int last = (int)t->n_levels - 1;
    while (sp) {
        --sp;
        const int64_t node = stack[sp];
        const int l = sl[sp];
        const double *b = t->box + 4 * (t->level_start[l] + node);
        if (!box_overlap(q, b)) continue;
        if (l == last) { vpush(hits, t->item_row[node]); continue; }
        const int64_t k1 = (node + 1) * t->fan < t->level_count[l + 1] ? (node + 1) * t->fan
                                                                        : t->level_count[l + 1];
        for (int64_t k = k1 - 1; k >= node * t->fan; --k) { stack[sp] = k; sl[sp] = l + 1; ++sp; }
    }
}

typedef struct {
    int64_t *left, *right; /* malloc'd, owned by caller after or_join_take */
    int64_t n;
} or_pairs;

static or_pairs g_last = {0, 0, 0};

/* predicate: 0 within (code == 2), 1 intersects (code != 0),
 *            2 covered_by (code != 0), 3 touches (code == 1).
 * Returns the number of pairs; fetch them with or_join_take. */
int64_t or_sjoin_points_polygons(const double *px, const double *py, int64_t n_points,
                                 const double *x, const double *y, const int32_t *geom,
                                 const int32_t *part, const int32_t *ring, int64_t n_polys,
                                 int predicate, int nthreads) {
    double *bounds = (double *)malloc(sizeof(double) * 4 * (n_polys ? n_polys : 1));
    or_polygon_bounds(x, y, geom, part, ring, n_polys, bounds);
    or_tree t;
    tree_build(&t, bounds, n_polys, 16);
    free(bounds);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const int nt = omp_get_max_threads();
#else
    const int nt = 1;
#endif
    i64vec *outs = (i64vec *)calloc((size_t)nt, sizeof(i64vec));
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
        const int nth = omp_get_num_threads();
#else
        const int tid = 0, nth = 1;
#endif
        const int64_t lo = n_points * tid / nth, hi = n_points * (tid + 1) / nth;
        i64vec hits = {0, 0, 0};
        i64vec *out = &outs[tid];
        for (int64_t i = lo; i < hi; ++i) {
            if (isnan(px[i]) || isnan(py[i])) continue;
            const double q[4] = {px[i], py[i], px[i], py[i]};
            hits.n = 0;
            tree_query(&t, q, &hits);
            if (hits.n > 1) qsort(hits.v, (size_t)hits.n, sizeof(int64_t), cmp_i64);
            for (int64_t h = 0; h < hits.n; ++h) {
                const int code = or_point_location(px[i], py[i], x, y, geom, part, ring, hits.v[h]);
                const int keep = predicate == 0 ? code == 2
                               : predicate == 3 ? code == 1
                                                : code != 0;
                if (keep) { vpush(out, i); vpush(out, hits.v[h]); }
            }
        }
        free(hits.v);
    }
    tree_free(&t);
    int64_t total = 0;
    for (int k = 0; k < nt; ++k) total += outs[k].n / 2;
    free(g_last.left); free(g_last.right);
    g_last.left = (int64_t *)malloc(sizeof(int64_t) * (total ? total : 1));
    g_last.right = (int64_t *)malloc(sizeof(int64_t) * (total ? total : 1));
    g_last.n = total;
    int64_t w = 0;
    for (int k = 0; k < nt; ++k) {
        for (int64_t j = 0; j < outs[k].n; j += 2) {
            g_last.left[w] = outs[k].v[j];
            g_last.right[w] = outs[k].v[j + 1];
            ++w;
        }
        free(outs[k].v);
    }
    free(outs);
    return total;
}

/* Polygon x polygon intersects join; left rows query a tree over right rows. */
int64_t or_sjoin_polygons_polygons(const double *lx, const double *ly, const int32_t *lg,
                                   const int32_t *lp, const int32_t *lr, int64_t nl,
                                   const double *rx, const double *ry, const int32_t *rg,
                                   const int32_t *rp, const int32_t *rr, int64_t nr,
                                   int nthreads) {
    double *lb = (double *)malloc(sizeof(double) * 4 * (nl ? nl : 1));
    double *rb = (double *)malloc(sizeof(double) * 4 * (nr ? nr : 1));
    or_polygon_bounds(lx, ly, lg, lp, lr, nl, lb);
    or_polygon_bounds(rx, ry, rg, rp, rr, nr, rb);
    or_tree t;
    tree_build(&t, rb, nr, 16);
    or_polys A = {lx, ly, lg, lp, lr}, B = {rx, ry, rg, rp, rr};
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const int nt = omp_get_max_threads();
#else
    const int nt = 1;
#endif
    i64vec *outs = (i64vec *)calloc((size_t)nt, sizeof(i64vec));
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
        const int nth = omp_get_num_threads();
#else
        const int tid = 0, nth = 1;
#endif
        const int64_t lo = nl * tid / nth, hi = nl * (tid + 1) / nth;
        i64vec hits = {0, 0, 0};
        for (int64_t i = lo; i < hi; ++i) {
            if (isnan(lb[4 * i])) continue;
            hits.n = 0;
            tree_query(&t, lb + 4 * i, &hits);
            if (hits.n > 1) qsort(hits.v, (size_t)hits.n, sizeof(int64_t), cmp_i64);
            for (int64_t h = 0; h < hits.n; ++h)
                if (poly_intersects(&A, i, &B, hits.v[h])) { vpush(&outs[tid], i); vpush(&outs[tid], hits.v[h]); }
        }
        free(hits.v);
    }
    tree_free(&t);
    free(lb); free(rb);
    int64_t total = 0;
    for (int k = 0; k < nt; ++k) total += outs[k].n / 2;
    free(g_last.left); free(g_last.right);
    g_last.left = (int64_t *)malloc(sizeof(int64_t) * (total ? total : 1));
    g_last.right = (int64_t *)malloc(sizeof(int64_t) * (total ? total : 1));
    g_last.n = total;
    int64_t w = 0;
    for (int k = 0; k < nt; ++k) {
        for (int64_t j = 0; j < outs[k].n; j += 2) {
            g_last.left[w] = outs[k].v[j]; g_last.right[w] = outs[k].v[j + 1]; ++w;
        }
        free(outs[k].v);
    }
    free(outs);
    return total;
}

/* Inclusive point-in-box query of boxes against a point set (sindex.query with
 * predicate='intersects' on a point tree); pairs are (box, point) lexsorted. */
int64_t or_box_query_points(const double *px, const double *py, int64_t n_points,
                            const double *boxes, int64_t n_boxes, int nthreads) {
    double *pb = (double *)malloc(sizeof(double) * 4 * (n_points ? n_points : 1));
    for (int64_t i = 0; i < n_points; ++i) {
        pb[4 * i] = pb[4 * i + 2] = px[i];
        pb[4 * i + 1] = pb[4 * i + 3] = py[i];
        if (isnan(px[i]) || isnan(py[i])) pb[4 * i] = pb[4 * i + 1] = pb[4 * i + 2] = pb[4 * i + 3] = NAN;
    }
    or_tree t;
    tree_build(&t, pb, n_points, 16);
    free(pb);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const int nt = omp_get_max_threads();
#else
    const int nt = 1;
#endif
    i64vec *outs = (i64vec *)calloc((size_t)nt, sizeof(i64vec));
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
        const int nth = omp_get_num_threads();
#else
        const int tid = 0, nth = 1;
#endif
        const int64_t lo = n_boxes * tid / nth, hi = n_boxes * (tid + 1) / nth;
        i64vec hits = {0, 0, 0};
        for (int64_t i = lo; i < hi; ++i) {
            if (isnan(boxes[4 * i])) continue;
            hits.n = 0;
            tree_query(&t, boxes + 4 * i, &hits);
            if (hits.n > 1) qsort(hits.v, (size_t)hits.n, sizeof(int64_t), cmp_i64);
            for (int64_t h = 0; h < hits.n; ++h) { vpush(&outs[tid], i); vpush(&outs[tid], hits.v[h]); }
        }
        free(hits.v);
    }
    tree_free(&t);
    int64_t total = 0;
    for (int k = 0; k < nt; ++k) total += outs[k].n / 2;
    free(g_last.left); free(g_last.right);
    g_last.left = (int64_t *)malloc(sizeof(int64_t) * (total ? total : 1));
    g_last.right = (int64_t *)malloc(sizeof(int64_t) * (total ? total : 1));
    g_last.n = total;
    int64_t w = 0;
    for (int k = 0; k < nt; ++k) {
        for (int64_t j = 0; j < outs[k].n; j += 2) {
            g_last.left[w] = outs[k].v[j]; g_last.right[w] = outs[k].v[j + 1]; ++w;
        }
        free(outs[k].v);
    }
    free(outs);
    return total;
}

/* Copy the last join's pairs out (n = value returned by the join). */
void or_join_take(int64_t *left, int64_t *right) {
    if (g_last.n) {
        memcpy(left, g_last.left, sizeof(int64_t) * g_last.n);
        memcpy(right, g_last.right, sizeof(int64_t) * g_last.n);
    }
    free(g_last.left); free(g_last.right);
    g_last.left = g_last.right = 0;
    g_last.n = 0;
}
