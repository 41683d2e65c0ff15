/*
 * oracle.c -- CPU restatement of the spatial-search hot path, used ONLY as test
 * infrastructure: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it as the checker.  Nothing in the product library links or calls it.
 *
 * Contracts restated here (reference file:line in each function's comment):
 *   - exact k-NN: ids ordered by ascending (d64, id), where
 *       d64 = ((dx*dx + dy*dy) + dz*dz), dx = (double)q.x - (double)p.x,
 *     every operation IEEE-754 binary64 round-to-nearest, no FMA contraction
 *     (this file must be built with -ffp-contract=off).
 *     Self-matches are included (gprt-macros/src/lib.rs:341 reports every point
 *     with d^2 <= worst; nothing filters d == 0 in the k_nn! register-heap path).
 *   - fixed-radius range query: all ids with d64 <= (double)r * (double)r,
 *     ordered by (d64, id)  (predicate of gprt-core/src/lib.rs:82-93, d^2 <= r^2).
 *   - 30-bit Morton codes (gprt-core/src/morton.rs:4-48).
 *   - LCG radius sampler (gprt-optix/src/dispatch.rs:65-103).
 *   - strided radius sampler of k_nn! (gprt-macros/src/lib.rs:417-450).
 *   - Barnes-Hut reference: FP64 direct sum of the Plummer-softened law
 *     a_i = sum_j G m_j (x_j - x_i) / (|x_j - x_i|^2 + eps^2)^{3/2}.
 *   - CPU exact k-NN path (KD-tree, median split, leaf 16, FP64, OpenMP) that is
 *     timed as the CPU baseline; it returns the same ids as the brute force.
 *
 * Pinning of each function is done in tests/test_oracle.py (closed forms,
 * scipy cKDTree cross-checks, committed golden fixtures, brute force on tiny
 * inputs, planted-mistake checks).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EXPORT __attribute__((visibility("default")))
#define ID_NONE 0xFFFFFFFFu

/* ------------------------------------------------------------------------ */
/* exact distance contract                                                   */
/* ------------------------------------------------------------------------ */
static inline double d2_64(const float *q, const float *p) {
    double dx = (double)q[0] - (double)p[0];
    double dy = (double)q[1] - (double)p[1];
    double dz = (double)q[2] - (double)p[2];
    return (dx * dx + dy * dy) + dz * dz;
}

EXPORT double oracle_d2_64(const float *q, const float *p) { return d2_64(q, p); }

/* (d, id) lexicographic "less than" */
static inline int lt_pair(double da, uint32_t ia, double db, uint32_t ib) {
    return da < db || (da == db && ia < ib);
}

/* insert (d,id) into an ascending list of capacity k holding *cnt entries */
static inline void topk_insert(double *dl, uint32_t *il, int k, int *cnt, double d, uint32_t id) {
    int c = *cnt;
    if (c == k) {
        if (!lt_pair(d, id, dl[k - 1], il[k - 1])) return;
        c = k - 1;
    }
    int j = c;
    while (j > 0 && lt_pair(d, id, dl[j - 1], il[j - 1])) {
        dl[j] = dl[j - 1];
        il[j] = il[j - 1];
        --j;
    }
    dl[j] = d;
    il[j] = id;
    if (*cnt < k) *cnt = *cnt + 1;
}

/* ------------------------------------------------------------------------ */
/* FP64 brute-force exact k-NN                                               */
/* ------------------------------------------------------------------------ */
/* Restates the result contract of k_nn! (gprt-macros/src/lib.rs:408-604):
 * k ids per query, ascending distance; ties broken by id (the reference leaves
 * them in AnyHit order, which is undefined).  When n < k the tail is padded
 * with 0xFFFFFFFF (the reference's empty-slot sentinel, lib.rs:317-320) and
 * +inf distances. */
EXPORT int oracle_knn_brute(const float *pts, int64_t n, const float *qs, int64_t nq, int k,
                            uint32_t *ids_out, double *d2_out, int nthreads) {
    if (k <= 0 || n < 0 || nq < 0) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
    {
        double *dl = (double *)malloc(sizeof(double) * (size_t)k);
        uint32_t *il = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)k);
#pragma omp for schedule(dynamic, 64)
        for (int64_t qi = 0; qi < nq; ++qi) {
            int cnt = 0;
            const float *q = qs + 3 * qi;
            for (int64_t i = 0; i < n; ++i)
                topk_insert(dl, il, k, &cnt, d2_64(q, pts + 3 * i), (uint32_t)i);
            for (int j = 0; j < k; ++j) {
                ids_out[qi * k + j] = j < cnt ? il[j] : ID_NONE;
                if (d2_out) d2_out[qi * k + j] = j < cnt ? dl[j] : INFINITY;
            }
        }
        free(dl);
        free(il);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FP64 brute-force range query (CSR output)                                 */
/* ------------------------------------------------------------------------ */
typedef struct { double d; uint32_t id; } pair_t;

static int cmp_pair(const void *a, const void *b) {
    const pair_t *x = (const pair_t *)a, *y = (const pair_t *)b;
    if (lt_pair(x->d, x->id, y->d, y->id)) return -1;
    if (lt_pair(y->d, y->id, x->d, x->id)) return 1;
    return 0;
}

/* Fixed-radius neighbours with d64 <= (double)r*(double)r, the d^2 <= r^2 test
 * of Sphere::intersect (gprt-core/src/lib.rs:82-93) evaluated exactly.
 * Phase 1 (ids == NULL): offsets[0..nq] = exclusive prefix of counts.
 * Phase 2: ids[offsets[q] .. offsets[q+1]) sorted by (d64, id). */
EXPORT int oracle_range_brute(const float *pts, int64_t n, const float *qs, int64_t nq, float r,
                              int64_t *offsets, uint32_t *ids, int nthreads) {
    double r2 = (double)r * (double)r;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    if (!ids) {
        offsets[0] = 0;
        int64_t *cnt = (int64_t *)calloc((size_t)nq + 1, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 64)
        for (int64_t qi = 0; qi < nq; ++qi) {
            int64_t c = 0;
            for (int64_t i = 0; i < n; ++i) c += d2_64(qs + 3 * qi, pts + 3 * i) <= r2;
            cnt[qi] = c;
        }
        for (int64_t qi = 0; qi < nq; ++qi) offsets[qi + 1] = offsets[qi] + cnt[qi];
        free(cnt);
        return 0;
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t qi = 0; qi < nq; ++qi) {
        int64_t b = offsets[qi], m = offsets[qi + 1] - b, c = 0;
        pair_t *buf = (pair_t *)malloc(sizeof(pair_t) * (size_t)(m > 0 ? m : 1));
        for (int64_t i = 0; i < n && c < m; ++i) {
            double d = d2_64(qs + 3 * qi, pts + 3 * i);
            if (d <= r2) { buf[c].d = d; buf[c].id = (uint32_t)i; ++c; }
        }
        qsort(buf, (size_t)c, sizeof(pair_t), cmp_pair);
        for (int64_t j = 0; j < c; ++j) ids[b + j] = buf[j].id;
        free(buf);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Morton codes (gprt-core/src/morton.rs:4-48)                               */
/* ------------------------------------------------------------------------ */
/* expand_bits, morton.rs:4-11 */
EXPORT uint32_t oracle_expand_bits(uint32_t x) {
    x &= 0x000003ffu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

/* morton, morton.rs:14-16 */
EXPORT uint32_t oracle_morton3(uint32_t x, uint32_t y, uint32_t z) {
    return (oracle_expand_bits(x) << 2) | (oracle_expand_bits(y) << 1) | oracle_expand_bits(z);
}

/* Rust `f32 as u32`: saturating, NaN -> 0 */
static inline uint32_t sat_u32(float f) {
    if (!(f > 0.0f)) return 0u; /* negative, zero, NaN */
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

/* compute_morton_codes, morton.rs:19-48: bbox over the points (f32 min/max),
 * extent = max(max - min, 1e-6), code = morton(min(sat((p-min)/ext*1023),1023)..) */
EXPORT int oracle_morton_codes(const float *pts, int64_t n, uint32_t *codes, float *bbox6) {
    if (n <= 0) return 0;
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            float v = pts[3 * i + a];
            mn[a] = fminf(mn[a], v);
            mx[a] = fmaxf(mx[a], v);
        }
    float ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = fmaxf(mx[a] - mn[a], 1e-6f);
    if (bbox6) { for (int a = 0; a < 3; ++a) { bbox6[a] = mn[a]; bbox6[3 + a] = mx[a]; } }
    for (int64_t i = 0; i < n; ++i) {
        uint32_t c[3];
        for (int a = 0; a < 3; ++a) {
            float t = (pts[3 * i + a] - mn[a]) / ext[a] * 1023.0f;
            uint32_t u = sat_u32(t);
            c[a] = u < 1023u ? u : 1023u;
        }
        codes[i] = oracle_morton3(c[0], c[1], c[2]);
    }
    return 0;
}

/* Same code for arbitrary points (e.g. queries) against a given bbox: the
 * per-point step of morton.rs:42-47 with the bbox passed in. */
EXPORT int oracle_morton_codes_bbox(const float *pts, int64_t n, const float *bbox6, uint32_t *codes) {
    float ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = fmaxf(bbox6[3 + a] - bbox6[a], 1e-6f);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t c[3];
        for (int a = 0; a < 3; ++a) {
            float t = (pts[3 * i + a] - bbox6[a]) / ext[a] * 1023.0f;
            uint32_t u = sat_u32(t);
            c[a] = u < 1023u ? u : 1023u;
        }
        codes[i] = oracle_morton3(c[0], c[1], c[2]);
    }
    return 0;
}

/* Wide keys of the tree build: the same quantisation with b bits per axis,
 * q = min(sat((p - min) / ext * (2^b - 1)), 2^b - 1), interleaved bit by bit
 * (x above y above z at each level).  At b = 10 this is morton.rs:42-47
 * exactly, which is how the tests pin it to the reference's codes. */
EXPORT int oracle_morton_codes_wide(const float *pts, int64_t n, const float *bbox6, int bits, uint64_t *codes) {
    if (bits < 1 || bits > 21) return 1;
    const uint32_t qmax = (1u << bits) - 1u;
    const float scale = (float)qmax;
    float ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = fmaxf(bbox6[3 + a] - bbox6[a], 1e-6f);
    for (int64_t i = 0; i < n; ++i) {
        uint32_t c[3];
        for (int a = 0; a < 3; ++a) {
            float t = (pts[3 * i + a] - bbox6[a]) / ext[a] * scale;
            uint32_t u = sat_u32(t);
            c[a] = u < qmax ? u : qmax;
        }
        uint64_t m = 0;
        for (int l = bits - 1; l >= 0; --l)
            for (int a = 0; a < 3; ++a) m = (m << 1) | ((c[a] >> l) & 1u);
        codes[i] = m;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* radius samplers                                                           */
/* ------------------------------------------------------------------------ */
/* heuristic: 0 = SampledMax, 1 = SampledPercentile(param), 2 = Fixed(param)
 * (RadiusHeuristic, gprt-ir/src/lib.rs:33-37) */

static int cmp_f32(const void *a, const void *b) {
    float x = *(const float *)a, y = *(const float *)b;
    return (x > y) - (x < y);
}

/* Rust `(x as f32 ...) as usize` for a non-negative finite float */
static inline int64_t sat_usize(float f) {
    if (!(f > 0.0f)) return 0;
    return (int64_t)f;
}

/* The LCG used by dispatch.rs:65-69: seed = seed*6364136223846793005 +
 * 1442695040888963407 (mod 2^64), draw = seed >> 33. */
EXPORT uint64_t oracle_lcg_draws(uint64_t seed, int64_t count, uint64_t *out) {
    for (int64_t i = 0; i < count; ++i) {
        seed = seed * 6364136223846793005ull + 1442695040888963407ull;
        if (out) out[i] = seed >> 33;
    }
    return seed;
}

/* LCG radius sampler, dispatch.rs:65-103.  100 query draws, each followed by
 * min(5000, n) point draws; keeps the 4 smallest sqrt(d^2) (k is ignored, no
 * self-exclusion) with the strict swap-insert of :82-87; pushes the 4th when
 * finite; sorts; picks by heuristic and multiplies by 1.1; 10.0 if no sample. */
EXPORT float oracle_sampler_lcg(const float *pts, int64_t n, const float *qs, int64_t nq, int heuristic,
                                float param) {
    uint64_t seed = 0x123456789ABCDEFull;
    int64_t sample_size = 5000 < n ? 5000 : n;
    int64_t qss = 100 < nq ? 100 : nq;
    float *kth = (float *)malloc(sizeof(float) * (size_t)(qss > 0 ? qss : 1));
    int64_t len = 0;
    for (int64_t s = 0; s < qss; ++s) {
        seed = seed * 6364136223846793005ull + 1442695040888963407ull;
        int64_t qi = (int64_t)((seed >> 33) % (uint64_t)nq);
        const float *q = qs + 3 * qi;
        float d[4] = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX};
        for (int64_t t = 0; t < sample_size; ++t) {
            seed = seed * 6364136223846793005ull + 1442695040888963407ull;
            int64_t pi = (int64_t)((seed >> 33) % (uint64_t)n);
            const float *p = pts + 3 * pi;
            float dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
            float dd = sqrtf((dx * dx + dy * dy) + dz * dz);
            if (dd < d[3]) {
                d[3] = dd;
                if (d[3] < d[2]) { float x = d[2]; d[2] = d[3]; d[3] = x; }
                if (d[2] < d[1]) { float x = d[1]; d[1] = d[2]; d[2] = x; }
                if (d[1] < d[0]) { float x = d[0]; d[0] = d[1]; d[1] = x; }
            }
        }
        if (d[3] != FLT_MAX) kth[len++] = d[3];
    }
    float radius = 10.0f;
    if (len > 0) {
        qsort(kth, (size_t)len, sizeof(float), cmp_f32);
        if (heuristic == 0) radius = kth[len - 1] * 1.1f;
        else if (heuristic == 1) {
            float a = (float)len * param, b = (float)(len - 1);
            int64_t idx = sat_usize(fminf(a, b));
            radius = kth[idx] * 1.1f;
        } else radius = param;
    }
    free(kth);
    return radius;
}

/* Strided sampler of k_nn!, gprt-macros/src/lib.rs:417-450.  Default radius
 * 1.0; Fixed(r) -> r; SampledPercentile(p): every max(nq/100,1)-th query (at
 * most 100) against every max(n/1000,1)-th point (at most 1000), keeping the k
 * smallest d^2 > 1e-10 (f32), pushing sqrt(k-th) when it is not f32::MAX,
 * index min((len*p) as usize, len-1), times 1.1.  SampledMax is unhandled
 * there (the `_ => {}` arm) and leaves 1.0. */
EXPORT float oracle_sampler_strided(const float *pts, int64_t n, const float *qs, int64_t nq, int k,
                                    int heuristic, float param) {
    float radius = 1.0f;
    if (heuristic == 2) return param;
    if (heuristic != 1 || k <= 0) return radius;
    int64_t qstep = nq / 100 > 1 ? nq / 100 : 1;
    int64_t pstep = n / 1000 > 1 ? n / 1000 : 1;
    float *dists = (float *)malloc(sizeof(float) * 100);
    float *knn = (float *)malloc(sizeof(float) * (size_t)k);
    int64_t len = 0;
    int64_t taken_q = 0;
    for (int64_t qi = 0; qi < nq && taken_q < 100; qi += qstep, ++taken_q) {
        const float *q = qs + 3 * qi;
        for (int j = 0; j < k; ++j) knn[j] = FLT_MAX;
        int64_t taken_p = 0;
        for (int64_t pi = 0; pi < n && taken_p < 1000; pi += pstep, ++taken_p) {
            const float *p = pts + 3 * pi;
            float dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
            float d2 = (dx * dx + dy * dy) + dz * dz;
            if (d2 > 1e-10f && d2 < knn[k - 1]) {
                knn[k - 1] = d2;
                qsort(knn, (size_t)k, sizeof(float), cmp_f32);
            }
        }
        if (knn[k - 1] != FLT_MAX) dists[len++] = sqrtf(knn[k - 1]);
    }
    if (len > 0) {
        qsort(dists, (size_t)len, sizeof(float), cmp_f32);
        int64_t idx = sat_usize((float)len * param);
        if (idx > len - 1) idx = len - 1;
        radius = dists[idx] * 1.1f;
    }
    free(dists);
    free(knn);
    return radius;
}

/* ------------------------------------------------------------------------ */
/* Barnes-Hut reference: FP64 direct sum                                     */
/* ------------------------------------------------------------------------ */
/* pos_mass: float4 (x, y, z, m) per body.  For each sampled body i:
 * acc[i] = sum_{j} G m_j (x_j - x_i) / (|x_j - x_i|^2 + eps^2)^{3/2}
 * (the j == i term is zero).  This is the Plummer-softened force law of the
 * Barnes-Hut workload; the reference's own law is oracle_bh_direct_ref. */
EXPORT int oracle_bh_direct(const float *pos_mass, int64_t n, const int64_t *sample, int64_t ns, double G,
                            double eps, double *acc_out, int nthreads) {
    double e2 = eps * eps;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t s = 0; s < ns; ++s) {
        int64_t i = sample ? sample[s] : s;
        double xi = pos_mass[4 * i], yi = pos_mass[4 * i + 1], zi = pos_mass[4 * i + 2];
        double ax = 0, ay = 0, az = 0;
        for (int64_t j = 0; j < n; ++j) {
            double dx = (double)pos_mass[4 * j] - xi;
            double dy = (double)pos_mass[4 * j + 1] - yi;
            double dz = (double)pos_mass[4 * j + 2] - zi;
            double r2 = (dx * dx + dy * dy) + dz * dz + e2;
            double inv = 1.0 / (r2 * sqrt(r2));
            double w = (double)pos_mass[4 * j + 3] * inv;
            ax += w * dx;
            ay += w * dy;
            az += w * dz;
        }
        acc_out[3 * s] = G * ax;
        acc_out[3 * s + 1] = G * ay;
        acc_out[3 * s + 2] = G * az;
    }
    return 0;
}

/* The reference's force law (gprt-macros/src/lib.rs:785, 790, 799), FP64: for
 * each sampled body i,
 * acc[i] = sum_{j : d2 > 1e-10} G m_j (x_j - x_i) / (d2 sqrt(d2 + eps^2)),
 * d2 = |x_j - x_i|^2 (the shader divides G m_i m_j by d2 sqrt(d2 + 1e-6), i.e.
 * eps = 1e-3, and skips pairs with d2 <= 1e-10, the body itself included).
 * As shipped, barnes_hut! opens every node (a null traversable), so its output
 * is this sum in FP32 times m_i. */
EXPORT int oracle_bh_direct_ref(const float *pos_mass, int64_t n, const int64_t *sample, int64_t ns, double G,
                                double eps, double *acc_out, int nthreads) {
    double e2 = eps * eps;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t s = 0; s < ns; ++s) {
        int64_t i = sample ? sample[s] : s;
        double xi = pos_mass[4 * i], yi = pos_mass[4 * i + 1], zi = pos_mass[4 * i + 2];
        double ax = 0, ay = 0, az = 0;
        for (int64_t j = 0; j < n; ++j) {
            double dx = (double)pos_mass[4 * j] - xi;
            double dy = (double)pos_mass[4 * j + 1] - yi;
            double dz = (double)pos_mass[4 * j + 2] - zi;
            double d2 = (dx * dx + dy * dy) + dz * dz;
            if (!(d2 > 1e-10)) continue;
            double w = (double)pos_mass[4 * j + 3] / (d2 * sqrt(d2 + e2));
            ax += w * dx;
            ay += w * dy;
            az += w * dz;
        }
        acc_out[3 * s] = G * ax;
        acc_out[3 * s + 1] = G * ay;
        acc_out[3 * s + 2] = G * az;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* CPU exact k-NN path: KD-tree, median split, leaf 16, FP64, OpenMP          */
/* ------------------------------------------------------------------------ */
typedef struct {
    float lo[3], hi[3];
    int64_t begin, end;  /* point range in perm */
    int64_t left, right; /* -1 for leaves */
} kd_node;

typedef struct {
    const float *pts;
    int64_t *perm;
    kd_node *nodes;
    int64_t nnodes;
} kd_tree;

static void kd_bbox(const float *pts, const int64_t *perm, int64_t b, int64_t e, float *lo, float *hi) {
    for (int a = 0; a < 3; ++a) { lo[a] = FLT_MAX; hi[a] = -FLT_MAX; }
    for (int64_t i = b; i < e; ++i)
        for (int a = 0; a < 3; ++a) {
            float v = pts[3 * perm[i] + a];
            if (v < lo[a]) lo[a] = v;
            if (v > hi[a]) hi[a] = v;
        }
}

/* nth_element on perm[b..e) by coordinate axis (quickselect) */
static void kd_select(const float *pts, int64_t *perm, int64_t b, int64_t e, int64_t m, int axis) {
    while (e - b > 1) {
        int64_t mid = b + (e - b) / 2;
        float pv = pts[3 * perm[mid] + axis];
        int64_t i = b, j = e - 1;
        while (i <= j) {
            while (pts[3 * perm[i] + axis] < pv) ++i;
            while (pts[3 * perm[j] + axis] > pv) --j;
            if (i <= j) { int64_t t = perm[i]; perm[i] = perm[j]; perm[j] = t; ++i; --j; }
        }
        if (m <= j) e = j + 1;
        else if (m >= i) b = i;
        else return;
    }
}

static int64_t kd_build_rec(kd_tree *t, int64_t b, int64_t e) {
    int64_t id = t->nnodes++;
    kd_node *nd = &t->nodes[id];
    kd_bbox(t->pts, t->perm, b, e, nd->lo, nd->hi);
    nd->begin = b;
    nd->end = e;
    nd->left = nd->right = -1;
    if (e - b <= 16) return id;
    int axis = 0;
    float ext = nd->hi[0] - nd->lo[0];
    for (int a = 1; a < 3; ++a)
        if (nd->hi[a] - nd->lo[a] > ext) { ext = nd->hi[a] - nd->lo[a]; axis = a; }
    int64_t m = b + (e - b) / 2;
    kd_select(t->pts, t->perm, b, e, m, axis);
    int64_t l = kd_build_rec(t, b, m);
    int64_t r = kd_build_rec(t, m, e);
    t->nodes[id].left = l;
    t->nodes[id].right = r;
    return id;
}

/* box distance with the same rounding steps as d2_64: monotone, so a box value
 * never exceeds d2_64 of any point inside the box */
static inline double box_d2_64(const float *q, const float *lo, const float *hi) {
    double c[3];
    for (int a = 0; a < 3; ++a) {
        float v = q[a] < lo[a] ? lo[a] : (q[a] > hi[a] ? hi[a] : q[a]);
        c[a] = (double)q[a] - (double)v;
    }
    return (c[0] * c[0] + c[1] * c[1]) + c[2] * c[2];
}

static void kd_query(const kd_tree *t, const float *q, int k, double *dl, uint32_t *il, int *cnt) {
    int64_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const kd_node *nd = &t->nodes[stack[--sp]];
        double bd = box_d2_64(q, nd->lo, nd->hi);
        if (*cnt == k && bd > dl[k - 1]) continue;
        if (nd->left < 0) {
            for (int64_t i = nd->begin; i < nd->end; ++i) {
                int64_t pid = t->perm[i];
                topk_insert(dl, il, k, cnt, d2_64(q, t->pts + 3 * pid), (uint32_t)pid);
            }
            continue;
        }
        const kd_node *L = &t->nodes[nd->left], *R = &t->nodes[nd->right];
        double dlb = box_d2_64(q, L->lo, L->hi), drb = box_d2_64(q, R->lo, R->hi);
        /* push far first so near is processed next */
        if (dlb <= drb) { stack[sp++] = nd->right; stack[sp++] = nd->left; }
        else { stack[sp++] = nd->left; stack[sp++] = nd->right; }
    }
}

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

EXPORT int oracle_kdtree_knn(const float *pts, int64_t n, const float *qs, int64_t nq, int k, uint32_t *ids_out,
                             double *d2_out, int nthreads, double *build_ms, double *query_ms) {
    if (k <= 0) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    double t0 = now_ms();
    kd_tree t;
    t.pts = pts;
    t.nnodes = 0;
    t.perm = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    t.nodes = (kd_node *)malloc(sizeof(kd_node) * (size_t)(2 * (n / 8 + 2)));
    for (int64_t i = 0; i < n; ++i) t.perm[i] = i;
    if (n > 0) kd_build_rec(&t, 0, n);
    double t1 = now_ms();
#pragma omp parallel
    {
        double *dl = (double *)malloc(sizeof(double) * (size_t)k);
        uint32_t *il = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)k);
#pragma omp for schedule(static)
        for (int64_t qi = 0; qi < nq; ++qi) {
            int cnt = 0;
            if (n > 0) kd_query(&t, qs + 3 * qi, k, dl, il, &cnt);
            for (int j = 0; j < k; ++j) {
                ids_out[qi * k + j] = j < cnt ? il[j] : ID_NONE;
                if (d2_out) d2_out[qi * k + j] = j < cnt ? dl[j] : INFINITY;
            }
        }
        free(dl);
        free(il);
    }
    double t2 = now_ms();
    if (build_ms) *build_ms = t1 - t0;
    if (query_ms) *query_ms = t2 - t1;
    free(t.perm);
    free(t.nodes);
    return 0;
}

EXPORT int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
