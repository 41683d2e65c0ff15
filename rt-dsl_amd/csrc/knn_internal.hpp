// knn_internal.hpp -- launch arguments and launchers of the k-NN kernels.
#pragma once
#include "gprt_internal.hpp"

namespace gprt {

// device counter slots (unsigned[CNT_TOTAL])
constexpr int MAX_ROUNDS = 16;
constexpr int CNT_ACTIVE = 0;     // [r]: queries active in round r (r = 0 .. MAX_ROUNDS)
constexpr int CNT_WORK = 20;      // [r]: work-queue head of round r
constexpr int CNT_SORTERR = 39;   // radix-sort look-back walks that gave up (must stay 0)
constexpr int CNT_FALLBACK = 40;  // queries sent to the exact kernel
constexpr int CNT_PARTIAL = 41;   // queries of clouds with fewer than k points
constexpr int CNT_ERROR = 42;     // traversals stopped by the step guard (corrupt tree)
constexpr int CNT_BADID = 43;     // re-ranked candidates outside their query's cloud (must stay 0)
constexpr int CNT_AUX = 44;       // [0..3]: scratch counters of the range pipeline (segment lists)
constexpr int CNT_OFFSETS = 48;   // range fill: queries whose hits do not fill their offsets range
constexpr int CNT_RETRY = 49;     // range fill: long segments left to the chunked merge sort
constexpr int CNT_TOTAL = 56;   // the last slot (CNT_TOTAL - 1) is the input-validation flag
constexpr int CNT_STATS64 = 6;   // u64 statistics after the counters: visits, tests, re-rank reads,
                                 // (spare), wave steps, lane steps
constexpr int CNT_ALL = CNT_TOTAL + 2 * CNT_STATS64;  // words of the counter buffer

// Karras depth bound: 30 Morton bits + up to 32 cloud bits... a path of the
// radix tree has strictly increasing common-prefix lengths, so its length is
// at most (key bits) + (index tie-break bits) = 30 + cloud bits + 32 <= 96.
constexpr int EXACT_STACK = 96;

struct RoundArgs {
    const Node* nodes;
    const Node4* nodes4;     // 4-wide view (wide per-lane walk)
    const float4* pts;
    const int2* roots;
    const float4* qs;        // queries in processing order, w = global query index
    const int* qcloud;       // cloud of each sorted query (nullptr: cloud 0)
    const int* kc;           // k per cloud (nullptr: uniform k)
    int k;
    const float* radius0;    // initial radius per cloud
    float mult;
    int round;
    int last;                // 1: unbounded round, every query resolves
    const int* in_list;      // nullptr: identity over [0, count)
    int* out_list;
    int2* fb_list;           // fallback: (sorted position, float bits of an upper bound on the k-th d64)
    unsigned* counters;
    unsigned long long* stats64;
    uint32_t* ids_out;
    double* d2_out;
    const int64_t* out_base; // per cloud output offset (nullptr: g * k)
    const int64_t* qoff;     // per cloud first global query index
    int max_steps;           // node-visit guard (> number of internal nodes)
    int base;                // first sorted position of this k class (identity work list)
    int rounds;              // rounds of this schedule (the last is unbounded)
    int inline_rounds;       // 1: an unresolved query retries with the next radius in the same launch
    int2* spill;             // stack entries below the LDS ring: [level][grid lanes] of (ref, d32)
    int spill_depth;         // levels of spill per lane
    const int64_t* poff;     // per cloud first sorted point (cloud_off of the tree): the id guard
    const float* q_xyz;      // wave-per-query plain form: the caller's xyz queries in input order
    int64_t nq_plain;        // and their number
};

struct ExactArgs {
    int mode;                // 0 plain (caller's xyz array, input order), 1 fallback list,
                             // 2 sorted positions [base, base + nq)
    int base;
    int64_t nq;
    const float* q_xyz;
    const float4* qs;
    const int2* fb_list;
    const int* qcloud;
    const int* kc;
    int k;
    const Node* nodes;
    const float4* pts;
    const int2* roots;
    unsigned* counters;
    unsigned long long* stats64;
    uint32_t* ids_out;
    double* d2_out;
    const int64_t* out_base;
    const int64_t* qoff;
    int max_steps;
};

int launch_sampler(hipStream_t s, int sampler, const float4* orig4, const int64_t* poff,
                   const float* qs,
                   const int64_t* qoff, int nclouds, int k, int heuristic, float param, float* kth, float* radius);
// queries -> Morton-bucket order (float4 with the input index in .w); keys is
// 8 B of scratch per query, scratch holds query_prep_scratch_bytes(); flag
// receives the input-validation result.  prefix_cleared: the words named by
// query_prep_clear_range were zeroed by the caller, ahead on the same stream.
int launch_query_prep(Ctx* ctx, const float* qs, const int64_t* qoff, int nclouds, const float* bbox,
                      int use_morton, int nq, uint64_t* keys, uint32_t* perm, float4* qs4, unsigned* counters,
                      int64_t max_per_cloud, int* qcloud, unsigned* scratch, unsigned* flag, const int* kc = nullptr,
                      int k_uniform = 1, int sort_classes = 0, bool prefix_cleared = false);
// the sort scratch words launch_query_prep would zero first (none when the
// order needs no sort); reserves that scratch
int query_prep_clear_range(Ctx* ctx, int nclouds, int use_morton, int nq, int sort_classes, unsigned** first,
                           int* words);
// k classes: queries of one class share a per-lane candidate list of
// capacity C with G = 3 guard slots, the smallest C of {8, 16, 24, 32, 40, 48}
// with k + G <= C (registers hold the list, so a tighter class runs at a
// higher occupancy).  Larger k (class QWAVE_CLASS) run one query per wavefront
// with the list spread over its lanes (knn_qwave.hip).
constexpr int N_KCLASS = 8;  // the query sort key holds the class in 3 bits
constexpr int N_LANE_CLASSES = 6;
constexpr int QWAVE_CLASS = 6;
__host__ __device__ inline int k_class_cap(int cls) {
    return cls == 0 ? 8 : (cls == 1 ? 16 : (cls == 2 ? 24 : (cls == 3 ? 32 : (cls == 4 ? 40 : 48))));
}
__host__ __device__ inline int k_class(int k) {
    int c = 0;
    while (c < N_LANE_CLASSES && k + 3 > k_class_cap(c)) ++c;
    return c;
}
size_t query_prep_scratch_bytes(int nclouds, int use_morton);
int launch_fill_u32(hipStream_t s, unsigned* p, int n, unsigned v);
// persistent grid (blocks) of the per-lane round for k class cls and an LDS
// ring of `ring` entries; grid_req > 0 is taken as given
int lane_round_grid(int cls, int ring, int grid_req, int trav);
// HBM spill levels per lane for a total stack depth, a k class and the
// configured LDS ring size
int lane_spill_depth(int stack_depth, int cls, int ring);
// per-file round launchers and resident grids (knn_lane.hip, knn_wide.hip,
// knn_packet.hip)
int lane_grid(int cls, int ring, int grid_req);
int wide_grid(int cls, int ring, int grid_req);
void launch_lane_round(hipStream_t s, const RoundArgs& a, int cls, int ring, int grid_req, bool count);
void launch_wide_round(hipStream_t s, const RoundArgs& a, int cls, int ring, int grid_req, bool count);
void launch_packet_round(hipStream_t s, const RoundArgs& a, int cls, int grid, bool count);
int launch_round(hipStream_t s, const RoundArgs& a, int cls, int grid, int trav, int ring, bool count);
int launch_exact(hipStream_t s, const ExactArgs& a, int kmax, int grid_threads, bool count);
// the wave-per-query kernel (knn_qwave.hip): over one k class of the batched
// form, or over the caller's queries in input order when a.q_xyz is set
int launch_qwave(hipStream_t s, const RoundArgs& a, int kmax, int grid_req, bool count);
int qwave_lanes_m(int kmax);
// the register-heap program for k > 64 (knn_qwave.hip)
int launch_heap_wave(hipStream_t s, const Node* nodes, const float4* pts, int2 root, const float4* q4, int nq, int k,
                     unsigned* heap, unsigned* lens, int max_steps, unsigned* err);

}  // namespace gprt
