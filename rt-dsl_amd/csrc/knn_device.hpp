// knn_device.hpp -- device-side pieces shared by the k-NN round kernels
// (knn_lane.hip, knn_wide.hip, knn_packet.hip): the fixed-slot FP32 candidate
// list, per-query state, the work queue, the FP64 re-rank and the leaf scan.
#pragma once
#include <atomic>

#include "gprt_internal.hpp"
#include "knn_internal.hpp"
#include "lane_walk.hpp"

namespace gprt {

// LDS ring sizes the per-lane kernels are built for
__host__ __device__ inline int ring_size(int ring) { return ring <= 8 ? 8 : (ring <= 16 ? 16 : 32); }

// LDS ring a k class runs with: class 0 (capacity 8) is built for 8, 16 and
// 32 entries; the larger classes, whose occupancy is bounded by registers, for 8
__host__ __device__ inline int effective_ring(int cls, int ring) { return cls == 0 ? ring_size(ring) : 8; }

// persistent grid of a round kernel: as many blocks of 256 as the device
// holds at once, cached per device in the caller's slot array
template <class F>
inline int resident_grid(F kernel, std::atomic<int>* cache) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::atomic<int>& slot = cache[dev & 63];
    int g = slot.load(std::memory_order_relaxed);
    if (g > 0) return g;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu <= 0) per_cu = 4;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    g = per_cu * cus;
    slot.store(g, std::memory_order_relaxed);
    return g;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// View of read-only device memory in the constant address space: wave-uniform
// loads through it become scalar loads (s_load_dwordx16 for a 64-byte node).
#define GPRT_AS4 __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const GPRT_AS4 T* const_view(const T* p) {
    return (const GPRT_AS4 T*)(uintptr_t)p;
}
__device__ __forceinline__ float4 ld4(const GPRT_AS4 float4* p) {
    const GPRT_AS4 float* f = (const GPRT_AS4 float*)p;
    return make_float4(f[0], f[1], f[2], f[3]);
}
__device__ __forceinline__ Node ldnode(const GPRT_AS4 Node* p) {
    const GPRT_AS4 float4* f = (const GPRT_AS4 float4*)p;
    Node n;
    n.alo = ld4(f);
    n.ahi = ld4(f + 1);
    n.blo = ld4(f + 2);
    n.bhi = ld4(f + 3);
    return n;
}
__device__ __forceinline__ int2 ld2i(const GPRT_AS4 int2* p) {
    const GPRT_AS4 int* f = (const GPRT_AS4 int*)p;
    return make_int2(f[0], f[1]);
}

// Adds the lanes that reach this point together (the active lanes of the
// wavefront) to *counter with one atomic: atomics of every lane on one counter
// serialise.
__device__ __forceinline__ void count_entering(unsigned* counter) {
    const uint64_t m = __ballot(true);
    if ((int)__lane_id() == __ffsll((unsigned long long)m) - 1) atomicAdd(counter, (unsigned)__popcll(m));
}

// Work queue of a persistent round: a wave takes the next 64 queries.  The
// head is read first, so once the queue is drained the remaining waves leave
// without contending on the atomic.
__device__ __forceinline__ unsigned grab_work(unsigned* head, unsigned n_in) {
    unsigned h = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (h >= n_in) return h;
    return atomicAdd(head, 64u);
}

// The same queue with the first task of every wavefront assigned by its index
// in the grid (tasks [0, waves) in the order the grid's waves are numbered),
// so the persistent grid's start does not queue all of its waves on one
// returning atomic (they serialise on the counter: 5,120 waves took about
// 60 us to get their first task); later tasks come from the counter, offset
// past those.  `first` tells which call this is for the wave.
__device__ __forceinline__ unsigned grab_work_first(unsigned* head, unsigned n_in, bool first, unsigned wave,
                                                    unsigned nwaves) {
    if (first) return wave * 64u;
    const unsigned skip = nwaves * 64u;
    if (skip >= n_in) return n_in;
    return skip + grab_work(head, n_in - skip);
}

// FP64 re-rank of the guard band in (d64, id) order and output of the first k.
// A candidate outside the query's cloud (which no correct walk produces) is
// counted in CNT_BADID, which fails the call, so a wrong id is never returned
// silently.
struct PtLoad {
    const float4* __restrict__ pts;
    __device__ __forceinline__ float4 operator()(int64_t i) const { return pts[i]; }
};
template <int C>
__device__ __forceinline__ int rerank_and_write(const RoundArgs& a, const float* ld, const int* ls, float qx,
                                                 float qy, float qz, float tf, int g, int c, int k,
                                                 const float4* __restrict__ pts) {
    double dd[C];
    uint32_t id[C];
    bool bad = false;
    const int reads = rerank_sort<C>(ld, ls, qx, qy, qz, tf, a.poff[c], a.poff[c + 1], PtLoad{pts}, dd, id, bad);
    if (bad) atomicAdd(&a.counters[CNT_BADID], 1u);
    const int64_t ob = a.out_base ? a.out_base[c] + (int64_t)(g - a.qoff[c]) * k : (int64_t)g * k;
#pragma unroll
    for (int j = 0; j < C; ++j)
        if (j < k) {
            a.ids_out[ob + j] = id[j];
            if (a.d2_out) a.d2_out[ob + j] = dd[j];
        }
    return reads;
}

// Scan a leaf bucket [f, f+m) issuing the point loads four at a time, so a
// bucket costs one memory latency per four points instead of one per point.
template <class Q4>
__device__ __forceinline__ void scan_bucket(Q4& Q, const float4* __restrict__ pts, int f, int m) {
    for (int t0 = 0; t0 < m; t0 += 4) {
        float4 p[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (t0 + u < m) p[u] = pts[f + t0 + u];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (t0 + u < m) Q.visit_point(p[u], f + t0 + u);
    }
}

}  // namespace gprt
