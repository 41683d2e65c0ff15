// knn_device.hpp -- device-side pieces shared by the k-NN round kernels
// (knn_lane.hip, knn_wide.hip, knn_packet.hip): the fixed-slot FP32 candidate
// list, per-query state, the work queue, the FP64 re-rank and the leaf scan.
#pragma once
#include <atomic>

#include "gprt_internal.hpp"
#include "knn_internal.hpp"

namespace gprt {

// spare candidate slots beyond k (the guard band)
constexpr int GUARD = 3;

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

// ---------------------------------------------------------------------------
// FP32 candidate list (ascending d32), capacity C, in registers.
//
// The list holds C = S + k + G slots: the first S = C - k - G are -inf
// sentinels that never move, then the k nearest candidates, then G guard slots
// for candidates inside the band of the k-th.  The k-th candidate therefore sits
// at the fixed index C - 1 - G whatever k is, so no slot is ever addressed with
// a runtime index (which would push the whole list to scratch memory).
// ---------------------------------------------------------------------------
template <int C>
struct CandList {
    float d[C];
    int s[C];
    __device__ __forceinline__ void init(int sentinels) {
#pragma unroll
        for (int j = 0; j < C; ++j) {
            d[j] = j < sentinels ? -INFINITY : INFINITY;
            s[j] = -1;
        }
    }
    // insert (dd, ss), dd >= 0; returns the value pushed out of the last slot
    // (INF if that slot was empty).  Caller guarantees dd < d[C-1].
    // The new slot j value of a sorted list is the median of the old d[j-1],
    // d[j] and dd (one v_med3_f32); ids follow with two selects.  An entry
    // equal to existing ones goes after them.
    __device__ __forceinline__ float insert(float dd, int ss) {
        const float ev = d[C - 1];
#pragma unroll
        for (int j = C - 1; j >= 1; --j) {
            const bool shift = d[j - 1] > dd;
            const bool here = d[j] > dd;
            s[j] = shift ? s[j - 1] : (here ? ss : s[j]);
            d[j] = __builtin_amdgcn_fmed3f(d[j - 1], d[j], dd);
        }
        s[0] = d[0] > dd ? ss : s[0];
        d[0] = fminf(d[0], dd);
        return ev;
    }
};

template <int C, int G>
struct F32Query {
    float qx, qy, qz;
    float T0;      // round bound (radius^2), INF in the unbounded round
    float T, tau;  // current k-th bound and guard threshold
    float ev_min;  // smallest d32 pushed out of a full list
    CandList<C> L;

    __device__ __forceinline__ void start(float x, float y, float z, float t0, int k, bool active) {
        qx = x; qy = y; qz = z;
        T0 = t0;
        T = t0;
        tau = active ? band(t0) : -1.0f;  // inactive lanes accept nothing
        ev_min = INFINITY;
        L.init(C - k - G);
    }
    // One branch per point (inside the band or not).  Within the band a point
    // that does not beat the last slot (rare: the guard slots are full) is
    // inserted as +inf, which leaves the sorted list unchanged, rather than
    // taking a second branch.
    __device__ __forceinline__ void visit_point(float4 p, int s) {
        const float dd = d2f(qx, qy, qz, p.x, p.y, p.z);
        if (dd <= tau) {
            const bool ins = dd < L.d[C - 1];
            const float ev = L.insert(ins ? dd : INFINITY, s);
            ev_min = fminf(ev_min, ins ? ev : dd);
            T = fminf(T0, L.d[C - 1 - G]);
            tau = band(T);
        }
    }
    __device__ __forceinline__ float kth() const { return L.d[C - 1 - G]; }
};

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

// Work queue of a persistent round: a wave takes the next 64 queries.  The
// head is read first, so once the queue is drained the remaining waves leave
// without contending on the atomic.
__device__ __forceinline__ unsigned grab_work(unsigned* head, unsigned n_in) {
    unsigned h = __hip_atomic_load(head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (h >= n_in) return h;
    return atomicAdd(head, 64u);
}

// FP64 re-rank of the guard band in (d64, id) order and output of the first k.
// Candidates are the list entries with 0 <= d32 <= tf (sentinels are -inf).
// Guard: every candidate's sorted position must lie in its cloud's range
// [poff[c], poff[c+1]) and its id below the cloud size; a violation (which no
// correct walk produces) is counted in CNT_BADID, which fails the call, so a
// wrong id is never returned silently.
template <int C>
__device__ __forceinline__ int rerank_and_write(const RoundArgs& a, const float* ld, const int* ls, float qx,
                                                 float qy, float qz, float tf, int g, int c, int k,
                                                 const float4* __restrict__ pts) {
    double dd[C];
    uint32_t id[C];
    int reads = 0;
    const int64_t lo = a.poff[c], hi = a.poff[c + 1];
    bool bad = false;
#pragma unroll
    for (int j = 0; j < C; ++j) {
        const bool in = ld[j] >= 0.0f && ld[j] <= tf && ls[j] >= 0;
        dd[j] = INFINITY;
        id[j] = GPRT_ID_NONE;
        reads += in;
        if (in) {
            const bool out = ls[j] < lo || ls[j] >= hi;
            const float4 p = pts[out ? lo : ls[j]];
            dd[j] = d2d(qx, qy, qz, p.x, p.y, p.z);
            id[j] = (uint32_t)__float_as_int(p.w);
            bad |= out || (int64_t)id[j] >= hi - lo;
        }
    }
    if (bad) atomicAdd(&a.counters[CNT_BADID], 1u);
    bool swapped = true;
    while (swapped) {
        swapped = false;
#pragma unroll
        for (int j = 0; j + 1 < C; ++j) {
            const bool gt = dd[j] > dd[j + 1] || (dd[j] == dd[j + 1] && id[j] > id[j + 1]);
            const double lo = gt ? dd[j + 1] : dd[j], hi = gt ? dd[j] : dd[j + 1];
            const uint32_t li = gt ? id[j + 1] : id[j], hiid = gt ? id[j] : id[j + 1];
            dd[j] = lo; dd[j + 1] = hi; id[j] = li; id[j + 1] = hiid;
            swapped |= gt;
        }
    }
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
