// knn_lane.hip -- one TrueKNN round (gprt-macros/src/lib.rs:466-594) as a
// persistent launch with one query per lane over the binary LBVH.
#include <algorithm>
#include <atomic>

#include "knn_device.hpp"

namespace gprt {

// Minimum wavefronts per SIMD the register allocation of each list capacity
// must allow (1: no constraint), for the k classes whose occupancy pays for
// the spills it costs (LiDAR, 400 K queries; unconstrained -> constrained):
//   C = 8:  5 -> 6 waves (85 -> 80 VGPRs, 7 spilled), k = 5 round kernel
//           0.244 -> 0.222 ms; 7 waves (17 spilled) is slower;
//   C = 16: 4 -> 5 waves (111 -> 96 VGPRs, 29 spilled), k = 10 0.48 -> 0.41 ms;
//   C = 40: 1 -> 2 waves (261 -> 256 registers, 6 spilled), k = 32 1.81 -> 1.24 ms.
// 6 x 1,024 SIMDs x 64 lanes hold 393 K queries, nearly a 400 K batch in one
// pass of the persistent grid.
#ifndef GPRT_LANE_MIN_WAVES
#define GPRT_LANE_MIN_WAVES 6
#endif
#ifndef GPRT_LANE_MIN_WAVES_16
#define GPRT_LANE_MIN_WAVES_16 5
#endif
#ifndef GPRT_LANE_MIN_WAVES_24
#define GPRT_LANE_MIN_WAVES_24 1
#endif
#ifndef GPRT_LANE_MIN_WAVES_32
#define GPRT_LANE_MIN_WAVES_32 1
#endif
#ifndef GPRT_LANE_MIN_WAVES_40
#define GPRT_LANE_MIN_WAVES_40 2
#endif
#ifndef GPRT_LANE_MIN_WAVES_48
#define GPRT_LANE_MIN_WAVES_48 1
#endif
template <int C>
constexpr int lane_min_waves() {
    return C == 8 ? GPRT_LANE_MIN_WAVES
                  : C == 16 ? GPRT_LANE_MIN_WAVES_16
                            : C == 24 ? GPRT_LANE_MIN_WAVES_24
                                      : C == 32 ? GPRT_LANE_MIN_WAVES_32
                                                : C == 40 ? GPRT_LANE_MIN_WAVES_40 : GPRT_LANE_MIN_WAVES_48;
}


// The lane's memory for lane_walk: nodes and points in HBM, its column of the
// block's LDS ring ([slot][lane], so lanes hit distinct banks) and its column
// of the HBM spill array ([level][grid lanes]).
#define GPRT_LDS __attribute__((address_space(3)))
template <int S>
struct LaneMemDev {
    const float4* __restrict__ nodes4;
    const float4* __restrict__ pts;
    GPRT_LDS int (*sref)[256];      // typed as LDS so every ring access is a ds_read / ds_write
    GPRT_LDS float (*sdist)[256];
    int tid;
    int2* __restrict__ spill;
    size_t gstride;
    // Trip vote: with GPRT_TRIP_VOTE a wavefront's trip is taken only by the
    // lanes of the kind (scanning a leaf range, or visiting a node) that at
    // least half of its walking lanes want, so each trip runs one of the two
    // paths instead of both; the others wait for a later trip.
    __device__ __forceinline__ bool turn(bool scanning) const {
#ifdef GPRT_TRIP_VOTE
        const uint64_t all = __ballot(true), sc = __ballot(scanning);
        const bool scan_turn = 2 * __popcll(sc) >= __popcll(all);
        return scanning == scan_turn;
#else
        (void)scanning;
        return true;
#endif
    }
    __device__ __forceinline__ void fetch4(bool scanning, int first, int node, int last, float4* v) const {
        const float4* src = scanning ? pts + first : nodes4 + 4 * (size_t)node;
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = src[min(u, last)];
    }
    __device__ __forceinline__ int ring_ref(int j) const { return sref[j][tid]; }
    __device__ __forceinline__ float ring_d(int j) const { return sdist[j][tid]; }
    __device__ __forceinline__ void ring_put(int j, int r, float d) const {
        sref[j][tid] = r;
        sdist[j][tid] = d;
    }
    __device__ __forceinline__ int2 spill_get(int level) const { return spill[level * gstride]; }
    __device__ __forceinline__ void spill_put(int level, int2 e) const { spill[level * gstride] = e; }
};

// ---------------------------------------------------------------------------
// The same round with one independent walk per lane (closest child first).
//
// Every trip of the walk loop issues exactly one 64-byte load per lane: either
// the next node (two child boxes) or the next four points of the leaf range
// being scanned.  A node's in-range leaf buckets (the two buckets of a node
// are adjacent ranges of the sorted points, so both form one range) are
// scanned over the following trips, and only then is the node's successor
// chosen, with the threshold those points left; so the walk visits exactly the
// nodes a scan-then-descend walk visits.  What the one-load trips buy is
// lockstep: a wavefront waits one memory latency per trip whether its lanes
// are descending or scanning, instead of one per node plus the longest scan
// of any lane.
//
// The stack holds (node ref, d32 of its box) pairs; a popped entry whose box
// is already beyond the threshold is dropped without loading the node.  The S
// most recent entries of each lane live in an LDS ring ([slot][lane], so lanes
// hit distinct banks); when the ring is full its oldest entry moves to the
// lane's HBM spill column, and is read back when the ring runs empty.  The
// bottom of a closest-first stack holds the far siblings of the first
// descent, which are mostly pruned by the time they are reached, so spills
// cost little while a small ring keeps occupancy high.  A walk deeper than
// S + spill_depth is finished by the exact FP64 kernel.
// ---------------------------------------------------------------------------
template <int C, int G, int S, bool COUNT, bool ROUND0>
__global__ __launch_bounds__(256, lane_min_waves<C>()) void knn_lane_kernel(RoundArgs a) {
    static_assert((S & (S - 1)) == 0, "LDS ring size must be a power of two");
    __shared__ int sref[S][256];
    __shared__ float sdist[S][256];
    const int tid = threadIdx.x, lane = tid & 63;
    const unsigned n_in = a.counters[CNT_ACTIVE + a.round];
    if (n_in == 0) return;
    const float4* __restrict__ pts = a.pts;
    const LaneMemDev<S> mem{(const float4*)a.nodes,
                            pts,
                            (GPRT_LDS int(*)[256])sref,
                            (GPRT_LDS float(*)[256])sdist,
                            tid,
                            a.spill + (size_t)blockIdx.x * 256 + tid,
                            (size_t)gridDim.x * 256};
    const int max_depth = S + a.spill_depth;
    uint64_t visits = 0, tests = 0, reranks = 0, lane_steps = 0, wave_steps = 0;
    const unsigned wave = (blockIdx.x * blockDim.x + tid) >> 6, nwaves = (gridDim.x * blockDim.x) >> 6;
    bool first = true;
    while (true) {
        unsigned base = 0;
        if (lane == 0) base = grab_work_first(&a.counters[CNT_WORK + a.round], n_in, first, wave, nwaves);
        base = __shfl(base, 0);
        first = false;
        if (base >= n_in) break;
        const unsigned i = base + lane;
        const bool valid = i < n_in;
        int task_steps = 0;
        int status = 0;
        float fbound = INFINITY;  // fallback: an upper bound on the k-th d64
        int s = 0;
        if (valid) {
            s = (!ROUND0 && a.in_list) ? a.in_list[i] : a.base + (int)i;
            const float4 q = a.qs[s];
            const int c = a.qcloud ? a.qcloud[s] : 0;
            const int k = a.kc ? a.kc[c] : a.k;
            // radius of this round; with inline rounds an unresolved query
            // retries here with the next radius instead of joining a list
            int round = a.round;
            float R = a.radius0[c];
            for (int r = 0; r < round; ++r) R *= a.mult;
            const int2 root = a.roots[c];
            F32Query<C, G> Q;
            bool overflow, resolved, last;
            float T0;
            while (true) {
            last = a.last || (a.inline_rounds && round == a.rounds - 1);
            T0 = last ? INFINITY : R * R;
            Q.start(q.x, q.y, q.z, T0, k, true);
            overflow = false;
            if (root.y > 0) task_steps += lane_walk<C, G, S, COUNT>(Q, root, mem, a.max_steps, max_depth, overflow,
                                                                   visits, tests);
            resolved = last || (Q.kth() < INFINITY && Q.kth() <= T0);
            if (resolved || overflow || !a.inline_rounds) break;
            ++round;
            R *= a.mult;
            count_entering(&a.counters[CNT_ACTIVE + round]);  // statistics: queries entering this round
            }
            const float kth = Q.kth();
            const bool full = kth < INFINITY;
            if (!resolved) {
                // next round's list; with inline rounds only a walk that outgrew
                // its stack is left unresolved, and it goes to the exact kernel
                status = a.inline_rounds ? 2 : 1;
                fbound = full ? band(kth) : INFINITY;
            } else {
                const float tf = band(kth);
                if (overflow || Q.ev_min <= tf) {
                    status = 2;
                    fbound = full ? band(kth) : INFINITY;
                } else {
                    const int rr = rerank_and_write<C>(a, Q.L.d, Q.L.s, Q.qx, Q.qy, Q.qz, tf, __float_as_int(q.w), c,
                                                       k, pts);
                    if (COUNT) reranks += rr;
                    if (!full) atomicAdd(&a.counters[CNT_PARTIAL], 1u);
                }
            }
        }
        if (COUNT) {
            int mx = task_steps;
            for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o));
            lane_steps += task_steps;
            if (lane == 0) wave_steps += mx;
        }
        const int slot = wave_append(&a.counters[CNT_ACTIVE + a.round + 1], status == 1);
        if (slot >= 0) a.out_list[slot] = s;
        const int fslot = wave_append(&a.counters[CNT_FALLBACK], status == 2);
        if (fslot >= 0) a.fb_list[fslot] = make_int2(s, __float_as_int(fbound));
    }
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) {
            visits += __shfl_xor(visits, o);
            tests += __shfl_xor(tests, o);
            reranks += __shfl_xor(reranks, o);
            lane_steps += __shfl_xor(lane_steps, o);
        }
        if (lane == 0) {
            atomicAdd(&a.stats64[0], (unsigned long long)visits);
            atomicAdd(&a.stats64[1], (unsigned long long)tests);
            atomicAdd(&a.stats64[2], (unsigned long long)reranks);
            atomicAdd(&a.stats64[4], (unsigned long long)wave_steps);
            atomicAdd(&a.stats64[5], (unsigned long long)lane_steps);
        }
    }
}


template <int C, int G, int S>
static int grid_cs(int grid_req) {
    if (grid_req > 0) return grid_req;
    static std::atomic<int> cache[64];
    return resident_grid(knn_lane_kernel<C, G, S, false, true>, cache);
}

int lane_grid(int cls, int ring, int grid_req) {
    switch (cls) {
        case 0:
            switch (ring_size(ring)) {
                case 8: return grid_cs<8, GUARD, 8>(grid_req);
                case 16: return grid_cs<8, GUARD, 16>(grid_req);
                default: return grid_cs<8, GUARD, 32>(grid_req);
            }
        case 1: return grid_cs<16, GUARD, 8>(grid_req);
        case 2: return grid_cs<24, GUARD, 8>(grid_req);
        case 3: return grid_cs<32, GUARD, 8>(grid_req);
        case 4: return grid_cs<40, GUARD, 8>(grid_req);
        default: return grid_cs<48, GUARD, 8>(grid_req);  // class 5 (k class 6 never runs here)
    }
}

// The first round (all queries, identity work list) gets its own kernel symbol
// so profiles report it apart from the short later rounds, which hold the few
// queries round 0 left unresolved (often none) and run on a quarter of the
// resident grid to keep their launches cheap.
template <int C, int G, int S>
static void launch_cs(hipStream_t s, const RoundArgs& a, int grid_req, bool count) {
    const int full = grid_cs<C, G, S>(grid_req);
    const int grid = (a.round == 0 || grid_req > 0) ? full : std::max(1, full / 4);
    if (a.round == 0) {
        if (count) hipLaunchKernelGGL((knn_lane_kernel<C, G, S, true, true>), dim3(grid), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((knn_lane_kernel<C, G, S, false, true>), dim3(grid), dim3(256), 0, s, a);
    } else {
        if (count) hipLaunchKernelGGL((knn_lane_kernel<C, G, S, true, false>), dim3(grid), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((knn_lane_kernel<C, G, S, false, false>), dim3(grid), dim3(256), 0, s, a);
    }
}

void launch_lane_round(hipStream_t s, const RoundArgs& a, int cls, int ring, int grid_req, bool count) {
    switch (cls) {
        case 0:
            switch (ring_size(ring)) {
                case 8: launch_cs<8, GUARD, 8>(s, a, grid_req, count); break;
                case 16: launch_cs<8, GUARD, 16>(s, a, grid_req, count); break;
                default: launch_cs<8, GUARD, 32>(s, a, grid_req, count); break;
            }
            break;
        case 1: launch_cs<16, GUARD, 8>(s, a, grid_req, count); break;
        case 2: launch_cs<24, GUARD, 8>(s, a, grid_req, count); break;
        case 3: launch_cs<32, GUARD, 8>(s, a, grid_req, count); break;
        case 4: launch_cs<40, GUARD, 8>(s, a, grid_req, count); break;
        default: launch_cs<48, GUARD, 8>(s, a, grid_req, count); break;  // class 5
    }
}

}  // namespace gprt
