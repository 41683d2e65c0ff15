/*
 * gprt_hip.h -- C ABI of libgprt_hip.so, the MI355X (gfx950) spatial-search
 * backend.  Plain pointers and sizes only.
 *
 * Two groups of entry points:
 *
 *  1. The 13 symbols of the reference's native FFI, same names and argument
 *     types, so the Rust wrapper crate can bind this library with its extern
 *     block unchanged (gprt-optix/src/lib.rs:13-31, implemented in
 *     gprt-optix/host/main.cpp:101-415).  They keep the reference's
 *     no-return-value signatures; failures are reported through
 *     gprt_hip_last_error() instead of exit().
 *
 *  2. gprt_hip_* entry points returning an int status (GPRT_OK == 0).  These
 *     carry the hot path: LBVH build, exact k-NN (plain FP64 form and the
 *     batched FP32 + FP64-recheck form), fixed-radius range query, Barnes-Hut
 *     forces, the radius samplers, Morton codes and run statistics.
 *
 * Result contract shared by every k-NN / range entry point:
 *   d64(q,p) = ((dx*dx + dy*dy) + dz*dz) in binary64 with dx = (double)q.x -
 *   (double)p.x, each operation rounded to nearest.  k-NN returns the k points
 *   with the smallest (d64, id) pairs in ascending order; self-matches are
 *   included (gprt-macros/src/lib.rs:341 has no d > 0 filter).  Missing slots
 *   (cloud smaller than k) hold 0xFFFFFFFF and +inf.  A range query returns all
 *   points with d64 <= (double)r * (double)r, ordered the same way.
 *
 * Pointer flags: GPRT_PTR_DEVICE means every array argument of the call is a
 * device pointer on the context's device (inputs already resident in HBM);
 * otherwise they are host pointers and the call copies through.
 *
 * Threading: a context owns one HIP stream and a mutex.  Every call that
 * uses a context, directly, through one of its trees or through a legacy
 * pipeline, holds that mutex until it returns, so calls on one context from
 * several threads are safe and run one after another (the Rust wrapper keeps
 * its caches behind a Mutex for the same reason, gprt-optix/src/lib.rs:35-40).
 * Calls on different contexts run concurrently.  Destroying a context or a
 * tree while another thread still uses it is an error.  Error messages are
 * per thread.
 */
#ifndef GPRT_HIP_H
#define GPRT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPRT_HIP_API __attribute__((visibility("default")))

/* ---------------- status codes ---------------- */
#define GPRT_OK 0
#define GPRT_ERR_ARG 1      /* invalid argument (null pointer, k <= 0, ...) */
#define GPRT_ERR_INPUT 2    /* NaN / Inf / |coordinate| > 1e18 in the input */
#define GPRT_ERR_HIP 3      /* HIP runtime failure */
#define GPRT_ERR_ALLOC 4    /* device allocation failed */
#define GPRT_ERR_CAPACITY 5 /* output capacity too small (range query) */
#define GPRT_ERR_STATE 6    /* object in the wrong state, or a device consistency guard tripped */

/* largest k of every k-NN entry point (the reference's k_nn! register heap of
 * 2k slots has no bound, gprt-macros/src/lib.rs:470-471) */
#define GPRT_HIP_K_MAX 4096

#define GPRT_PTR_HOST 0
#define GPRT_PTR_DEVICE 1
/* with GPRT_PTR_DEVICE: only enqueue on the context stream and return; the
 * caller synchronises with gprt_hip_ctx_sync() (statistics are then partial) */
#define GPRT_NO_SYNC 2
/* gprt_hip_knn_plain, gprt_hip_range, gprt_hip_bh_forces: count node visits
 * and point tests (Barnes-Hut: also far-field interactions) into the
 * statistics (slower: for the roofline figures) */
#define GPRT_COLLECT_COUNTERS 4
/* tree build: order the points by the reference's 30-bit Morton codes (10 bits
 * per axis, gprt-core/src/morton.rs:4-48) instead of the default 21 bits per
 * axis.  Results are the same either way; only the tree shape changes. */
#define GPRT_KEYS_REF30 8
/* gprt_hip_bh_forces: the reference shader's force law
 * G m_j d / (|d|^2 sqrt(|d|^2 + eps^2)) over pairs with |d|^2 > 1e-10
 * (gprt-macros/src/lib.rs:785; there eps = 1e-3) instead of Plummer softening */
#define GPRT_BH_REFERENCE_LAW 16
/* gprt_hip_range: time the traversal kernel of the call (count walk in the
 * first phase, fill walk in the second) alone, into ms_kernel */
#define GPRT_TIME_KERNELS 32
/* tree build: split the Morton-ordered points top-down by binned SAH (8 bins
 * per axis, stable partitions) instead of the Karras radix hierarchy; same
 * node layout, so every query runs on either tree with the same results.
 * Fewer node visits on surface-like data, at a higher build cost.  No
 * counterpart in the reference (its GAS build is OptiX's,
 * gprt-optix/host/main.cpp:200-248). */
#define GPRT_BUILD_SAH 64
/* gprt_hip_bh_forces: walk the tree one body per lane, each lane on its own
 * path, instead of the default walk of 64 Morton-adjacent bodies per
 * wavefront over the union of their paths with wave-uniform node loads (each
 * body keeps its own opening decisions there too).  Both give the same bits;
 * the per-lane walk is kept for comparison. */
#define GPRT_BH_LANE_WALK 128

#define GPRT_ID_NONE 0xFFFFFFFFu

/* RadiusHeuristic of gprt-ir/src/lib.rs:33-37, plus INFINITE (one unbounded
 * round, no sampler) */
#define GPRT_RADIUS_MAX 0
#define GPRT_RADIUS_PERCENTILE 1
#define GPRT_RADIUS_FIXED 2
#define GPRT_RADIUS_INFINITE 3

/* initial-radius sampler */
#define GPRT_SAMPLER_LCG 0     /* gprt-optix/src/dispatch.rs:65-103 */
#define GPRT_SAMPLER_STRIDED 1 /* gprt-macros/src/lib.rs:417-450 */

/* traversal mapping */
#define GPRT_TRAV_THREAD 0 /* one query per lane */
#define GPRT_TRAV_WAVE 1   /* 64 Morton-adjacent queries walk the tree together (packet) */
#define GPRT_TRAV_WIDE 2   /* one query per lane over the 4-wide view of the tree */

typedef struct gprt_hip_ctx gprt_hip_ctx;
typedef struct gprt_hip_tree gprt_hip_tree;

/* Tunable knobs of the batched k-NN form.  Superset of the reference's
 * Schedule (gprt-ir/src/lib.rs:58-89): radius_heuristic/radius_param map to
 * RadiusHeuristic, radius_mult to radius_increment_mult, sort_queries to
 * use_morton_lbv.  max_hits_per_query and memory_strategy have no role here:
 * candidates live in registers and the output is k ids per query. */
typedef struct {
    int sampler;          /* GPRT_SAMPLER_*, default LCG */
    int radius_heuristic; /* GPRT_RADIUS_*, default PERCENTILE */
    float radius_param;   /* percentile p (default 0.95) or the fixed radius */
    float radius_mult;    /* radius growth per round (default 3.0) */
    int max_rounds;       /* bounded rounds before the final unbounded one (default 2) */
    int traversal;        /* GPRT_TRAV_*, default THREAD */
    int stack_depth;      /* traversal-stack entries per query, LDS + HBM spill (default 64);
                             a deeper walk is finished by the exact FP64 kernel */
    int guard_slots;      /* spare candidate slots beyond k: must be 3 (the per-lane lists carry
                             exactly 3; other values are rejected); kept in the schedule format */
    int sort_queries;     /* process queries in Morton order (default 1) */
    int use_graph;        /* replay the pipeline as a captured hipGraph (default 1) */
    int grid_blocks;      /* persistent grid size, 0 = automatic */
    int block_threads;    /* 0 = automatic (256) */
    int collect_counters; /* count node visits / point tests (slower; for the roofline) */
    int time_kernels;     /* with use_graph == 0: time the first traversal round on its own */
    int lds_stack;        /* most recent stack entries kept in LDS per query: 8, 16 or 32
                             (default 8); older ones spill to an HBM workspace */
    int compact_rounds;   /* 1: one launch per round over the compacted list of unresolved queries
                             (the reference's round structure); 0 (default): an unresolved query
                             retries with the next radius inside the first round's launch */
} gprt_hip_config;

typedef struct {
    int rounds_run;            /* bounded rounds that had work */
    int64_t queries;           /* queries in the last call */
    int64_t fallback_queries;  /* queries re-ranked by the exact FP64 kernel */
    int64_t partial_queries;   /* queries whose cloud holds fewer than k points */
    int64_t active_per_round[16];
    float initial_radius;      /* sampled radius of cloud 0 */
    double ms_total;           /* device time of the last call (hipEvents) */
    double ms_query;           /* query phase only (resident -> resident) */
    double ms_build;           /* last tree build */
    int64_t node_visits;       /* only with counters enabled (Barnes-Hut: nodes opened) */
    int64_t point_tests;       /* (Barnes-Hut: body-body interactions) */
    int64_t rerank_reads;      /* FP64 re-rank point reads (counters enabled; Barnes-Hut: far-field
                                  node interactions; range query: output slots its walk writes, one
                                  per hit found by a point test and two per subtree taken whole) */
    double ms_kernel;          /* first traversal round alone (time_kernels; range: GPRT_TIME_KERNELS) */
    int64_t wave_steps;        /* per-lane rounds: sum over wavefront tasks of their longest walk in
                                  loop iterations (counters enabled); lane steps / (64 * this) is the
                                  SIMD efficiency of the walk loop.  Barnes-Hut (wavefront walk):
                                  nodes loaded, summed over wavefronts */
    int64_t lane_steps;        /* per-lane rounds: loop iterations summed over lanes (counters enabled).
                                  Barnes-Hut (wavefront walk): bucket bodies loaded, summed over
                                  wavefronts */
} gprt_hip_run_stats;

typedef struct {
    int64_t n;
    int nclouds;
    int leaf_size;
    float bbox[6];  /* cloud 0: min xyz, max xyz */
    int key_bits;   /* significant bits of the sort key: 3 * axis_bits + cloud bits */
    int axis_bits;  /* Morton bits per axis; key = cloud << (3 * axis_bits) | morton */
} gprt_hip_tree_desc;

/* ---------------- library / context ---------------- */
GPRT_HIP_API int gprt_hip_version(void);
GPRT_HIP_API const char* gprt_hip_last_error(void);
/* Empty this thread's last-error message, so that a caller of the
 * no-return-value reference symbols can tell whether a call failed. */
GPRT_HIP_API void gprt_hip_clear_error(void);
GPRT_HIP_API void gprt_hip_config_default(gprt_hip_config* cfg);
GPRT_HIP_API int gprt_hip_ctx_create(int device, gprt_hip_ctx** out);
/* Trees built on the context may outlive it: they stay usable with another
 * context on the same device (gprt_hip_tree_export excepted) and free their
 * own buffers when destroyed. */
GPRT_HIP_API int gprt_hip_ctx_destroy(gprt_hip_ctx* ctx);
GPRT_HIP_API int gprt_hip_ctx_sync(gprt_hip_ctx* ctx);
GPRT_HIP_API int gprt_hip_stats(gprt_hip_ctx* ctx, gprt_hip_run_stats* out);
GPRT_HIP_API int gprt_hip_num_devices(int* out);

/* Device buffers on the context's device, for callers that keep inputs and
 * outputs resident in HBM (GPRT_PTR_DEVICE).  kind: 0 host->device,
 * 1 device->host, 2 device->device.  Copies are ordered on the context stream
 * and complete before the call returns. */
GPRT_HIP_API int gprt_hip_device_alloc(gprt_hip_ctx* ctx, size_t bytes, void** out);
GPRT_HIP_API int gprt_hip_device_free(gprt_hip_ctx* ctx, void* ptr);
GPRT_HIP_API int gprt_hip_copy(gprt_hip_ctx* ctx, void* dst, const void* src, size_t bytes, int kind);

/* ---------------- LBVH ---------------- */
/* Build a GPU LBVH over n points (xyz float32, 12 bytes each): 30-bit Morton
 * codes normalised per axis as gprt-core/src/morton.rs:19-48, stable LSD radix
 * sort, Karras hierarchy, atomic bottom-up refit.  Replaces the GAS build of
 * gprt_bvh_build (gprt-optix/host/main.cpp:200-248) and Scene::build
 * (gprt-core/src/lib.rs:135-150).  leaf_size in [1, 32]. */
GPRT_HIP_API int gprt_hip_tree_build(gprt_hip_ctx* ctx, const float* pts, int64_t n, int leaf_size, int ptr_flags,
                                     gprt_hip_tree** out);
/* Many independent clouds in one tree: cloud c owns points
 * [cloud_offsets[c], cloud_offsets[c+1]) of pts (host array of nclouds+1).
 * Each cloud gets its own Morton normalisation and its own subtree. */
GPRT_HIP_API int gprt_hip_tree_build_batched(gprt_hip_ctx* ctx, const float* pts, const int64_t* cloud_offsets,
                                             int nclouds, int leaf_size, int ptr_flags, gprt_hip_tree** out);
/* Waits for the device, then hands the tree's device buffers to the context
 * that built it, whose next build reuses them (growing any that are too
 * small); a context holds at most one destroyed tree's buffers, freed by
 * gprt_hip_ctx_destroy. */
GPRT_HIP_API int gprt_hip_tree_destroy(gprt_hip_tree* tree);
GPRT_HIP_API int gprt_hip_tree_info(const gprt_hip_tree* tree, gprt_hip_tree_desc* out);
/* Host copies of the tree for invariant checks: sorted keys (n), permutation
 * (n), nodes (16 floats per internal node, max(n-1,0) of them), roots (2 ints
 * per cloud).  Any pointer may be NULL. */
GPRT_HIP_API int gprt_hip_tree_export(gprt_hip_tree* tree, uint64_t* keys, uint32_t* perm, float* nodes,
                                      int32_t* roots);

/* ---------------- exact k-NN ---------------- */
/* Batched FP32 + FP64-recheck form (the k_nn! driver of
 * gprt-macros/src/lib.rs:408-604 with TrueKNN rounds on the device): sampled
 * initial radius, Morton-ordered queries, FP32 traversal keeping a guard band
 * of candidates, FP64 re-rank by (d64, id), on-device compaction of unresolved
 * queries between rounds, queries whose band overflows re-ranked by the exact
 * kernel.  k = 1..GPRT_HIP_K_MAX; k > 45 runs one query per wavefront in
 * binary64 with the candidate list spread over the wavefront's lanes.
 * ids_out: nq*k; d2_out (may be NULL): nq*k binary64 d64 values. */
GPRT_HIP_API int gprt_hip_knn(gprt_hip_ctx* ctx, gprt_hip_tree* tree, const float* queries, int64_t nq, int k,
                              const gprt_hip_config* cfg, uint32_t* ids_out, double* d2_out, int ptr_flags);
/* Plain single-launch form: one query per lane in input order, FP64 distance
 * arithmetic throughout, initial radius +inf, no compaction, no graph.
 * k = 1..GPRT_HIP_K_MAX; k > 64 runs one query per wavefront. */
GPRT_HIP_API int gprt_hip_knn_plain(gprt_hip_ctx* ctx, gprt_hip_tree* tree, const float* queries, int64_t nq,
                                    int k, uint32_t* ids_out, double* d2_out, int ptr_flags);
/* Uneven batches: query set c (rows [query_offsets[c], query_offsets[c+1]))
 * searches cloud c of a batched tree with its own k = k_per_cloud[c] (1..GPRT_HIP_K_MAX).
 * Results of query set c start at out_offsets[c] = sum_{c'<c} nq_c' * k_c'.
 * query_offsets, k_per_cloud: host arrays. */
GPRT_HIP_API int gprt_hip_knn_batched(gprt_hip_ctx* ctx, gprt_hip_tree* tree, const float* queries,
                                      const int64_t* query_offsets, const int* k_per_cloud,
                                      const gprt_hip_config* cfg, uint32_t* ids_out, double* d2_out,
                                      int ptr_flags);

/* Allocation-free variant of gprt_hip_knn, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Queries, ids_out
 * and d2_out (may be NULL) are device pointers; the tree is built beforehand.
 * Every per-call workspace lives in one 32 MiB device array per GPU, held
 * until the process ends; the traversal stack is the LDS ring only (deeper
 * walks finish in the exact kernel) and no graph is captured.  Same results as
 * gprt_hip_knn.  nq above gprt_hip_knn_noalloc_max_queries() is rejected with
 * GPRT_ERR_CAPACITY.  All calls on one GPU share that array: calls that overlap
 * on one GPU (from several contexts or threads of different contexts) race.
 * Synchronous. */
GPRT_HIP_API int gprt_hip_knn_noalloc(gprt_hip_ctx* ctx, gprt_hip_tree* tree, const float* queries, int64_t nq,
                                      int k, const gprt_hip_config* cfg, uint32_t* ids_out, double* d2_out);
GPRT_HIP_API int64_t gprt_hip_knn_noalloc_max_queries(void);

/* ---------------- fixed-radius range query ---------------- */
/* r_nn! / range_query! (gprt-macros/src/lib.rs:729-750, which as shipped
 * returns one flat id list without per-query split).  Phase 1: ids_out ==
 * NULL fills offsets_out[0..nq] (exclusive prefix of per-query counts).
 * Phase 2: ids_out with capacity ids_capacity >= offsets_out[nq] receives the
 * ids of query q at [offsets_out[q], offsets_out[q+1]), ordered by (d64, id).
 * offsets_out and ids_out follow ptr_flags.  Phase 2 checks every query's
 * range against the hits it finds: offsets that do not belong to these
 * queries and this radius give GPRT_ERR_ARG (nothing is written outside
 * [0, ids_capacity)). */
GPRT_HIP_API int gprt_hip_range(gprt_hip_ctx* ctx, gprt_hip_tree* tree, const float* queries, int64_t nq,
                                float radius, int64_t* offsets_out, uint32_t* ids_out, int64_t ids_capacity,
                                int ptr_flags);

/* ---------------- Barnes-Hut ---------------- */
/* Accelerations a_i = sum_j G m_j (x_j - x_i) / (|x_j - x_i|^2 + eps^2)^1.5
 * with the Barnes-Hut approximation over an LBVH of the bodies: a node of
 * size s (largest box edge) at distance d from x_i to its centre of mass is
 * taken whole when s < theta * d.  FP32 traversal, compensated (Neumaier)
 * accumulation.  Force law: Plummer, G m_j d / (|d|^2 + eps^2)^1.5, or with
 * GPRT_BH_REFERENCE_LAW in ptr_flags the reference's; at theta = 0 the latter
 * is what barnes_hut! returns as shipped (every node opened) divided by m_i.  pos_mass: float4 (x, y, z, m); acc_out: 3 floats per body.
 * leaf_size: bodies per leaf bucket (1..32).  A bucket that is opened is
 * summed body by body, so larger buckets trade time for accuracy: on a
 * 1M-body Plummer sphere at theta 0.5, leaf 8 gives a median relative error
 * of 1.1e-3 in 10.3 ms, leaf 1 gives 1.3e-3 in 10.2 ms.  The Plummer law needs
 * eps > 0 (GPRT_ERR_ARG otherwise: every body's own term would be 0/0); the
 * reference law takes eps = 0.
 * Replaces barnes_hut! (gprt-macros/src/lib.rs:752-834). */
GPRT_HIP_API int gprt_hip_bh_forces(gprt_hip_ctx* ctx, const float* pos_mass, int64_t n, float theta, float eps,
                                    float G, int leaf_size, float* acc_out, int ptr_flags);

/* ---------------- pieces exposed for parity tests ---------------- */
/* 30-bit Morton codes of n points with the bbox of those points
 * (gprt-core/src/morton.rs:19-48).  bbox_out: 6 floats (may be NULL). */
GPRT_HIP_API int gprt_hip_morton_codes(gprt_hip_ctx* ctx, const float* pts, int64_t n, uint32_t* codes_out,
                                       float* bbox_out, int ptr_flags);
/* Initial radius of the given sampler / heuristic (k is used by the strided
 * sampler only).  Runs on the device. */
GPRT_HIP_API int gprt_hip_sample_radius(gprt_hip_ctx* ctx, const float* pts, int64_t n, const float* queries,
                                        int64_t nq, int sampler, int heuristic, float param, int k,
                                        float* radius_out, int ptr_flags);
/* Stable LSD radix sort of (key, value) pairs by bits [begin_bit, end_bit). */
GPRT_HIP_API int gprt_hip_sort_pairs(gprt_hip_ctx* ctx, uint64_t* keys, uint32_t* vals, int64_t n, int begin_bit,
                                     int end_bit, int ptr_flags);

/* ---------------- autotuner ---------------- */
/* Times candidate configurations (traversal mapping, stack depth, leaf size,
 * radius percentile and multiplier) on the given data and queries and writes
 * the fastest to cfg_out and *leaf_size_out; with json_path != NULL also
 * writes it as a Schedule-superset JSON file.  Replaces gprt_autotune!
 * (gprt-macros/src/lib.rs:120-214) and the funnel profiler
 * (gprt-profiler/src/main.rs:37-133). */
GPRT_HIP_API int gprt_hip_autotune(gprt_hip_ctx* ctx, const float* pts, int64_t n, const float* queries, int64_t nq,
                                   int k, int budget_trials, gprt_hip_config* cfg_out, int* leaf_size_out,
                                   const char* json_path, int ptr_flags);
/* As gprt_hip_autotune, also choosing the tree: the LBVH or the SAH build
 * (the Schedule's build_strategy, gprt-ir/src/lib.rs:46-49, 64), returned as
 * the build flags to pass to gprt_hip_tree_build (GPRT_BUILD_SAH or 0) and
 * written to the JSON file as build_strategy PreferFastTrace / PreferFastBuild.
 * With build_flags_out NULL it is gprt_hip_autotune (LBVH only). */
GPRT_HIP_API int gprt_hip_autotune_ex(gprt_hip_ctx* ctx, const float* pts, int64_t n, const float* queries,
                                      int64_t nq, int k, int budget_trials, gprt_hip_config* cfg_out,
                                      int* leaf_size_out, int* build_flags_out, const char* json_path, int ptr_flags);
/* The autotuner for uneven batches: the same search over the batched call of
 * gprt_hip_knn_batched on a tree of the given clouds (gprt_hip_tree_build_batched),
 * every candidate's ids equal to the default schedule's over all clouds.
 * cloud_offsets, query_offsets, k_per_cloud: host arrays as in those calls;
 * pts and queries host or device by ptr_flags.  build_flags_out may be NULL
 * (LBVH only).  Replaces gprt_autotune! (gprt-macros/src/lib.rs:120-214) for
 * the batched workload, which the reference tunes one data set at a time. */
GPRT_HIP_API int gprt_hip_autotune_batched(gprt_hip_ctx* ctx, const float* pts, const int64_t* cloud_offsets,
                                           int nclouds, const float* queries, const int64_t* query_offsets,
                                           const int* k_per_cloud, int budget_trials, gprt_hip_config* cfg_out,
                                           int* leaf_size_out, int* build_flags_out, const char* json_path,
                                           int ptr_flags);

/* ---------------- testing hooks ---------------- */
/* Morton bits (8..30, default 16) by which the batched k-NN orders its queries
 * before the rounds.  Process-wide and not synchronised: set it only while no
 * other thread calls the library.  Results never depend on it; a captured
 * graph keeps the ordering it was captured with (use_graph = 0 to compare). */
GPRT_HIP_API void gprt_hip_debug_set_query_sort_bits(int bits);
/* Device allocations the library made for its own buffers since load (calls),
 * and those still held (buffers, bytes); any pointer may be NULL.  No
 * counterpart in the reference (which exits on any failure, main.cpp:12-28):
 * for leak checks. */
GPRT_HIP_API void gprt_hip_debug_alloc_stats(int64_t* calls, int64_t* live_buffers, int64_t* live_bytes);
/* Fault injection for tests: the next `after` device allocations of the
 * library succeed, every later one fails with GPRT_ERR_ALLOC, until called
 * again; after < 0 turns it off.  The environment variable GPRT_HIP_FAULT=alloc
 * (or alloc:N) sets it at load.  Process-wide. */
GPRT_HIP_API void gprt_hip_debug_fail_alloc(int64_t after);
/* A plain streaming copy kernel (stream_copy_kernel) of `bytes` (a multiple of
 * 16) between device buffers: reads and writes exactly that many bytes, the
 * reference point for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings.  No
 * counterpart in the reference.  Synchronous. */
GPRT_HIP_API int gprt_hip_debug_stream_copy(gprt_hip_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---------------- sandbox inputs ---------------- */
#define GPRT_CSV_POINTS 0 /* "x,y,z[,...]" per line (gprt-sandbox/src/io.rs:8-30) */
#define GPRT_CSV_BODIES 1 /* "x,y,z,m[,...]" per line (gprt-sandbox/src/bin/barneshut.rs:8-34) */

/* Read a text point file as the reference's sandbox does.  POINTS: a line
 * whose first three ','-separated fields (trimmed) are not all numbers, or
 * that is not UTF-8, is skipped.  BODIES: blank lines and lines with fewer
 * than four fields are skipped, a field that is not a number reads as 0, a
 * body with mass <= 0 is skipped, and a line that is not UTF-8 fails with
 * GPRT_ERR_INPUT.  Rows keep the file's order (row i is id i); numbers are
 * rounded once to binary32.  Writes 3 (POINTS) or 4 (BODIES) floats per row
 * to out and the row count to *rows; when the file holds more than capacity
 * rows, writes nothing, sets *rows and returns GPRT_ERR_ARG (capacity 0
 * counts).  Replaces load_points_fast (io.rs:8) and load_galaxy
 * (barneshut.rs:8).  Host only; parses on all host cores. */
GPRT_HIP_API int gprt_hip_read_points(const char* path, int format, float* out, int64_t capacity, int64_t* rows);

/* Read a Schedule / wisdom JSON file's text (gprt-ir/src/lib.rs:58-89, the
 * *_wisdom.json files) into a config.  Missing fields take the reference's
 * Schedule::default() (lib.rs:76-87), so files without build_strategy /
 * geom_type load; a "gprt_hip" object (as gprt_hip_autotune writes) sets the
 * config fields it names and *leaf_size (default 8).  Host only. */
GPRT_HIP_API int gprt_hip_schedule_from_json(const char* json, size_t size, gprt_hip_config* cfg, int* leaf_size);
/* As gprt_hip_schedule_from_json, also reading build_strategy into the tree
 * build flags: PreferFastTrace (the reference's default, gprt-ir/src/lib.rs:85)
 * gives GPRT_BUILD_SAH, PreferFastBuild 0.  build_flags may be NULL. */
GPRT_HIP_API int gprt_hip_schedule_from_json_ex(const char* json, size_t size, gprt_hip_config* cfg, int* leaf_size,
                                                int* build_flags);

/* ---------------- device programs from the shader IR ---------------- */
/* Lower a program of the reference's shader IR (ShaderNode trees,
 * gprt-ir/src/lib.rs:8-18, serialised as JSON; the forms are listed at the top
 * of codegen.hip) to a HIP translation unit, as gprt-codegen's compile_program
 * does to CUDA (gprt-codegen/src/lib.rs:78-182).  Two-phase: with src_out NULL
 * only *src_size (bytes incl. the terminating NUL) is set.  Host only. */
GPRT_HIP_API int gprt_hip_codegen(const char* ir_json, size_t ir_size, char* src_out, size_t src_capacity,
                                  size_t* src_size);
/* Lower and compile with hiprtc for `arch` (NULL: gfx950), replacing the nvcc
 * step of gprt-codegen/src/lib.rs:184-203; *code_size receives the size of the
 * code object.  Host only (no GPU needed); errors carry the compiler log. */
GPRT_HIP_API int gprt_hip_codegen_compile(const char* ir_json, size_t ir_size, const char* arch, size_t* code_size);

/* ---------------- the reference FFI (gprt-optix/src/lib.rs:13-31) ---------------- */
typedef struct CGprtPipeline CGprtPipeline;
typedef struct CGprtBvh CGprtBvh;

/* main.cpp:101-180.  ir_data selects the device program: NULL/empty or a
 * descriptor "gprt-hip:knn-heap" (register-heap k-NN, default),
 * "gprt-hip:push" (per-query push of every hit; an array registered for
 * one query but executed with several rays collects all their hits in one
 * list, as gprt_trace! does), "gprt-hip:autorope" (the
 * Barnes-Hut walk), or "gprt-hip:ir:" followed by a shader-IR program in JSON,
 * compiled here for the GPU (see gprt_hip_codegen; NULL on a codegen or
 * compile error, with the message in gprt_hip_last_error).  Any other bytes
 * (an OptiX-IR blob) select the default. */
GPRT_HIP_API CGprtPipeline* gprt_pipeline_create(const void* ir_data, size_t ir_size, int is_hw_triangle);
GPRT_HIP_API void gprt_pipeline_destroy(CGprtPipeline* pipe);
/* main.cpp:200-248: h_geom holds float4 spheres (x, y, z, r); the LBVH is
 * built over the centres and the sphere radius is recorded. */
GPRT_HIP_API CGprtBvh* gprt_bvh_build(CGprtPipeline* pipe, const float* h_geom, const float* h_aabbs, int count,
                                      int geom_bytes);
/* main.cpp:250-252: builds over the triangle AABB centres. */
GPRT_HIP_API CGprtBvh* gprt_bvh_build_triangles(CGprtPipeline* pipe, const float* h_verts,
                                                const unsigned int* h_indices, int num_triangles,
                                                const float* h_aabbs);
/* main.cpp:254-286: records the new sphere radius; no rebuild is needed. */
GPRT_HIP_API void gprt_bvh_refit(CGprtPipeline* pipe, CGprtBvh* bvh, float radius, int count);
GPRT_HIP_API void gprt_bvh_destroy(CGprtBvh* bvh);
/* main.cpp:296-305 */
GPRT_HIP_API void gprt_register_array(CGprtPipeline* pipe, const char* name, unsigned int capacity_per_query,
                                      int num_queries);
/* main.cpp:307-311 */
GPRT_HIP_API void gprt_register_value(CGprtPipeline* pipe, const char* name);
/* main.cpp:313-341: h_queries holds float4 (x, y, z, radius). */
GPRT_HIP_API void gprt_execute(CGprtPipeline* pipe, CGprtBvh* bvh, const float* h_queries, int count);
/* main.cpp:343-353 */
GPRT_HIP_API void gprt_retrieve_array_lengths(CGprtPipeline* pipe, const char* name, unsigned int* h_lengths,
                                              int num_queries);
/* main.cpp:355-365 */
GPRT_HIP_API void gprt_retrieve_array_flat(CGprtPipeline* pipe, const char* name, unsigned int* h_out,
                                           int total_elements);
/* main.cpp:367-369 */
GPRT_HIP_API void gprt_retrieve_value(CGprtPipeline* pipe, const char* name, unsigned int* h_out);
/* main.cpp:371-415: walks the caller's flattened octree (Body 28 B,
 * NodeSpatial 16 B, NodeRouting 32 B) with its autorope links. */
GPRT_HIP_API void gprt_execute_autorope_soa_ffi(CGprtPipeline* pipe, const void* h_bodies, int num_bodies,
                                                const void* h_spatial, int num_nodes, const void* h_routing,
                                                float theta, void* h_out_forces);

#ifdef __cplusplus
}
#endif
#endif /* GPRT_HIP_H */
