// knn_qwave.hip -- exact k-NN for large k (k + 3 > 48, up to GPRT_HIP_K_MAX):
// one query per wavefront, binary64 distances throughout, and a candidate list
// spread over the wavefront's 64 lanes.
//
// The reference's register heap holds 2k slots per query in global memory and
// inserts each AnyHit candidate by a bubble pass over them
// (gprt-macros/src/lib.rs:317-320, 347-377, 470-471), so it takes any k.  A
// per-lane list in registers cannot follow it past a few dozen slots: the list
// and its FP64 re-rank take about five registers per slot, and one wavefront
// per SIMD is all that fits at 64 slots.  Here the 64 lanes of a wavefront
// share one query instead:
//   * the list is 64*M slots of (d64, id) in ascending order, lane l holding
//     slots [l*M, l*M + M); an insert is one compare per slot plus one
//     cross-lane move of the previous lane's last slot (DPP wave_shr), so its
//     cost grows with M = ceil(k / 64), not with k;
//   * the walk is wave-uniform: scalar node loads, closest child first, and a
//     stack of (node, box d64 lower bound) held two entries per lane, written
//     by a lane select and read with v_readlane;
//   * a child holding at most QW_SCAN (128) points is not descended into: its
//     point range (children of a Karras node cover adjacent ranges of the
//     sorted points) is tested 64 points at a time, one per lane, and the lanes
//     whose (d64, id) beats the current k-th are inserted in lane order (or,
//     with one slot per lane and 8 or more of them, merged at once).
// Every comparison is on (d64, id) in binary64, so the ids are exact by
// construction: there is no guard band and no fallback.  A box whose d64 equals
// the k-th is still entered (it may hold a smaller id).  Rounds follow the
// per-lane kernels: a query resolves in a bounded round when its k-th d64 is at
// most R^2, else it retries at once with the next radius; the last round is
// unbounded.
#include <algorithm>
#include <atomic>

#include "knn_device.hpp"

namespace gprt {

constexpr int QW_STACK = 128;  // two (ref, d) entries per lane; the Karras depth bound is <= 96
#ifndef GPRT_QW_SCAN
#define GPRT_QW_SCAN 128
#endif
constexpr int QW_SCAN = GPRT_QW_SCAN;  // children with at most this many points are scanned, not descended into

__device__ __forceinline__ bool lex_lt(double a, uint32_t ia, double b, uint32_t ib) {
    return a < b || (a == b && ia < ib);
}

// lane l receives lane l-1's value, lane 0 receives `first` (DPP wave_shr:1,
// a GFX9 data-parallel move; bound_ctrl off keeps `first` in lane 0)
__device__ __forceinline__ int from_prev_lane(int v, int first) {
    return __builtin_amdgcn_update_dpp(first, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ double from_prev_lane(double v, double first) {
    const int lo = from_prev_lane((int)__double2loint(v), (int)__double2loint(first));
    const int hi = from_prev_lane((int)__double2hiint(v), (int)__double2hiint(first));
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_d(double v, int l) {
    const int lo = __builtin_amdgcn_readlane((int)__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane((int)__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// 64*M slots of (d64, id), ascending; the first 64*M - k slots hold (-inf, 0)
// sentinels that never move, so the k-th entry is always the last slot (lane
// 63, position M - 1) whatever k is.
template <int M>
struct WaveList {
    double d[M];
    uint32_t id[M];
    // live slots start at (live_d, live_id): only a candidate below that pair enters
    __device__ __forceinline__ void init(int k, int lane, double live_d = INFINITY, uint32_t live_id = GPRT_ID_NONE) {
        const int s0 = 64 * M - k;
#pragma unroll
        for (int p = 0; p < M; ++p) {
            const bool sent = lane * M + p < s0;
            d[p] = sent ? -INFINITY : live_d;
            id[p] = sent ? 0u : live_id;
        }
    }
    __device__ __forceinline__ void worst(double& wd, uint32_t& wi) const {
        wd = readlane_d(d[M - 1], 63);
        wi = (uint32_t)__builtin_amdgcn_readlane((int)id[M - 1], 63);
    }
    // insert the wave-uniform candidate (dd, ii), known to beat the last slot.
    // Slot j's new value is slot j-1's if that one is greater than the
    // candidate, the candidate if slot j is greater, else its own.
    __device__ __forceinline__ void insert(double dd, uint32_t ii) {
        bool g[M];
#pragma unroll
        for (int p = 0; p < M; ++p) g[p] = lex_lt(dd, ii, d[p], id[p]);
        const double pd = from_prev_lane(d[M - 1], -INFINITY);
        const uint32_t pi = (uint32_t)from_prev_lane((int)id[M - 1], 0);
        const bool gp = lex_lt(dd, ii, pd, pi);
#pragma unroll
        for (int p = M - 1; p >= 0; --p) {
            const bool sh = p ? g[p - 1] : gp;
            const double prd = p ? d[p - 1] : pd;
            const uint32_t pri = p ? id[p - 1] : pi;
            d[p] = sh ? prd : (g[p] ? dd : d[p]);
            id[p] = sh ? pri : (g[p] ? ii : id[p]);
        }
    }
};

// value of the lane `lane ^ J`: a DPP quad permutation for J = 1, 2 (no LDS
// traffic), ds_swizzle in its bitmask mode (and 0x1f, xor J) within 32 lanes
// for J = 4, 8, 16, ds_bpermute for J = 32
template <int J>
__device__ __forceinline__ int xor_lane_i(int v) {
    if constexpr (J == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xf, 0xf, false);  // quad_perm [1,0,3,2]
    else if constexpr (J == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xf, 0xf, false);  // quad_perm [2,3,0,1]
    else if constexpr (J < 32) return __builtin_amdgcn_ds_swizzle(v, 0x1F | (J << 10));
    else return __shfl_xor(v, J);
}

// (d64, id) of the lane `lane ^ J`
template <int J>
__device__ __forceinline__ void xor_lane(double d, uint32_t i, double& od, uint32_t& oi) {
    od = __hiloint2double(xor_lane_i<J>((int)__double2hiint(d)), xor_lane_i<J>((int)__double2loint(d)));
    oi = (uint32_t)xor_lane_i<J>((int)i);
}

// One compare-exchange stage of a bitonic network over the 64 lanes: the lane
// keeps the smaller of its pair's two (d64, id) keys when keep_min is set,
// else the larger.  Keys are distinct pairs except for equal sentinels, whose
// exchange changes nothing.
template <int J>
__device__ __forceinline__ void bitonic_stage(double& d, uint32_t& i, bool keep_min) {
    double od;
    uint32_t oi;
    xor_lane<J>(d, i, od, oi);
    const bool other_lt = lex_lt(od, oi, d, i);
    if (other_lt == keep_min) {
        d = od;
        i = oi;
    }
}

// Merges up to 64 candidates (one per lane; (+inf, NONE) where a lane has
// none) into a one-slot-per-lane list (M = 1): the candidates are sorted into
// descending order by a bitonic network, the lane-wise minimum of the
// ascending list and the descending candidates is the bitonic sequence of the
// 64 smallest keys of both (the half-cleaner of Batcher's merge), and a
// bitonic merge puts it in ascending order.  The result equals inserting the
// candidates one by one.  About 27 stages of 3 lane exchanges each, which is
// less than sequential inserts from about 8 candidates on.
// one sort step of the candidates: pairs (l, l ^ J) in blocks of S lanes; pairs
// in blocks with (lane & S) set ascend, the others descend, so the final block
// (S = 64) descends
template <int S, int J>
__device__ __forceinline__ void sort_step(double& cd, uint32_t& ci, int lane) {
    const bool asc = (lane & S) != 0, lower = (lane & J) == 0;
    bitonic_stage<J>(cd, ci, lower == asc);
}
template <int J>
__device__ __forceinline__ void merge_step(double& ld, uint32_t& li, int lane) {
    bitonic_stage<J>(ld, li, (lane & J) == 0);
}

__device__ __forceinline__ void merge_batch(double& ld, uint32_t& li, double cd, uint32_t ci, int lane) {
    sort_step<2, 1>(cd, ci, lane);
    sort_step<4, 2>(cd, ci, lane);
    sort_step<4, 1>(cd, ci, lane);
    sort_step<8, 4>(cd, ci, lane);
    sort_step<8, 2>(cd, ci, lane);
    sort_step<8, 1>(cd, ci, lane);
    sort_step<16, 8>(cd, ci, lane);
    sort_step<16, 4>(cd, ci, lane);
    sort_step<16, 2>(cd, ci, lane);
    sort_step<16, 1>(cd, ci, lane);
    sort_step<32, 16>(cd, ci, lane);
    sort_step<32, 8>(cd, ci, lane);
    sort_step<32, 4>(cd, ci, lane);
    sort_step<32, 2>(cd, ci, lane);
    sort_step<32, 1>(cd, ci, lane);
    sort_step<64, 32>(cd, ci, lane);
    sort_step<64, 16>(cd, ci, lane);
    sort_step<64, 8>(cd, ci, lane);
    sort_step<64, 4>(cd, ci, lane);
    sort_step<64, 2>(cd, ci, lane);
    sort_step<64, 1>(cd, ci, lane);
    if (lex_lt(cd, ci, ld, li)) {
        ld = cd;
        li = ci;
    }
    merge_step<32>(ld, li, lane);
    merge_step<16>(ld, li, lane);
    merge_step<8>(ld, li, lane);
    merge_step<4>(ld, li, lane);
    merge_step<2>(ld, li, lane);
    merge_step<1>(ld, li, lane);
}

#ifndef GPRT_QW_BATCH
#define GPRT_QW_BATCH 8
#endif
constexpr int QW_BATCH = GPRT_QW_BATCH;  // candidates of one scan from which M = 1 lists merge them as a batch
constexpr unsigned QW_GRAB = 4;          // queries a wavefront takes from the work counter at a time

// one query's walk state
template <int M>
struct WaveQuery {
    float qx, qy, qz;
    double T0;          // round bound (R^2), +inf in the unbounded round
    double wd;          // current k-th (d64, id), the pruning bound together with T0
    uint32_t wi;
    WaveList<M> L;
    uint64_t tests;

    __device__ __forceinline__ double bound() const { return fmin(wd, T0); }

    // test the sorted points [f, f + m), one per lane
    __device__ __forceinline__ void scan(const float4* __restrict__ pts, int f, int m, int lane) {
        for (int t0 = 0; t0 < m; t0 += 64) {
            const int t = t0 + lane;
            const bool v = t < m;
            const float4 p = pts[f + (v ? t : m - 1)];
            const double dd = d2d(qx, qy, qz, p.x, p.y, p.z);
            const uint32_t ii = (uint32_t)__float_as_int(p.w);
            const bool cand = v && dd <= T0 && lex_lt(dd, ii, wd, wi);
            uint64_t mask = __ballot(cand);
            if (M == 1 && __popcll((unsigned long long)mask) >= QW_BATCH) {
                merge_batch(L.d[0], L.id[0], cand ? dd : INFINITY, cand ? ii : GPRT_ID_NONE, lane);
                L.worst(wd, wi);
                mask = 0;
            }
            while (mask) {
                const int j = __builtin_amdgcn_readfirstlane(__ffsll((unsigned long long)mask) - 1);
                mask &= mask - 1;
                const double cd = readlane_d(dd, j);
                const uint32_t ci = (uint32_t)__builtin_amdgcn_readlane((int)ii, j);
                if (lex_lt(cd, ci, wd, wi)) {
                    L.insert(cd, ci);
                    L.worst(wd, wi);
                }
            }
        }
        tests += (uint64_t)m;
    }
};

// first sorted point under a child: a leaf bucket starts at ~ref; an internal
// Karras node sits at the upper end of its range when it is child A and at the
// lower end when it is child B
__device__ __forceinline__ int child_first(int ref, int cnt, bool is_b) {
    return ref < 0 ? ~ref : (is_b ? ref : ref - cnt + 1);
}

// The whole walk of one query; false when the step guard trips (corrupt tree).
template <int M>
__device__ __forceinline__ bool qwave_walk(WaveQuery<M>& Q, const GPRT_AS4 Node* nodes, const float4* __restrict__ pts,
                                           int2 root, int lane, int max_steps, uint64_t& visits) {
    if (root.y <= 0) return true;
    if (root.x < 0) {
        Q.scan(pts, ~root.x, root.y, lane);
        return true;
    }
    int sref0 = 0, sref1 = 0;   // stack entry e lives in lane e & 63 of half e >> 6
    float sd0 = 0.f, sd1 = 0.f;
    int sp = 0, node = uniform(root.x), steps = 0;
    while (true) {
        if (++steps > max_steps) return false;
        const Node nd = ldnode(nodes + node);
        ++visits;
        const double dA = box2d(Q.qx, Q.qy, Q.qz, nd.alo, nd.ahi);
        const double dB = box2d(Q.qx, Q.qy, Q.qz, nd.blo, nd.bhi);
        const int rA = uniform(__float_as_int(nd.alo.w)), rB = uniform(__float_as_int(nd.blo.w));
        const int cA = uniform(__float_as_int(nd.ahi.w)), cB = uniform(__float_as_int(nd.bhi.w));
        double w = Q.bound();
        const bool inA = uniform(dA <= w), inB = uniform(dB <= w);
        const bool scA = inA && cA <= QW_SCAN, scB = inB && cB <= QW_SCAN;
        if (scA && scB) {
            Q.scan(pts, child_first(rA, cA, false), cA + cB, lane);  // adjacent ranges
        } else if (scA || scB) {
            if (scA) Q.scan(pts, child_first(rA, cA, false), cA, lane);
            else Q.scan(pts, child_first(rB, cB, true), cB, lane);
        }
        w = Q.bound();
        const bool gA = inA && !scA && uniform(dA <= w), gB = inB && !scB && uniform(dB <= w);
        if (gA && gB) {
            if (sp >= QW_STACK) return false;
            const bool bfirst = uniform(dB < dA);
            const int far = bfirst ? rA : rB;
            const float fd = f64_lower_f32(bfirst ? dA : dB);
            const bool mine = lane == (sp & 63);
            if (sp < 64) {
                sref0 = mine ? far : sref0;
                sd0 = mine ? fd : sd0;
            } else {
                sref1 = mine ? far : sref1;
                sd1 = mine ? fd : sd1;
            }
            ++sp;
            node = bfirst ? rB : rA;
            continue;
        }
        if (gA) { node = rA; continue; }
        if (gB) { node = rB; continue; }
        node = -1;
        while (sp > 0) {
            --sp;
            const int r = __builtin_amdgcn_readlane(sp < 64 ? sref0 : sref1, sp & 63);
            const float d = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sp < 64 ? sd0 : sd1), sp & 63));
            if ((double)d <= w) { node = r; break; }
        }
        if (node < 0) return true;
    }
}

// Persistent launch over the queries of one k class (batched form: sorted
// positions a.base + i, i < the class size in a.counters[CNT_ACTIVE]) or over
// the caller's xyz array in input order (plain form: a.q_xyz set, a.last set).
#ifndef GPRT_QW_MIN_WAVES
#define GPRT_QW_MIN_WAVES 1
#endif
template <int M, bool COUNT>
__global__ __launch_bounds__(256, M == 1 ? GPRT_QW_MIN_WAVES : 1) void knn_qwave_kernel(RoundArgs a) {
    const int lane = threadIdx.x & 63;
    const bool plain = a.q_xyz != nullptr;
    const unsigned n_in = plain ? (unsigned)a.nq_plain : a.counters[CNT_ACTIVE + a.round];
    const GPRT_AS4 Node* nodes = const_view(a.nodes);
    const unsigned wave0 = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nwaves = (gridDim.x * blockDim.x) >> 6;
    uint64_t visits = 0, tests = 0;
    unsigned qi = wave0;
    // batched form: queries are taken QW_GRAB at a time from the work counter
    // (one atomic per chunk: atomics of every wavefront on one counter
    // serialise), the first chunk of each wavefront by its index in the grid
    unsigned q_next = wave0 * QW_GRAB, q_end = q_next + QW_GRAB;
    const unsigned skip = nwaves * QW_GRAB;
    while (true) {
        if (!plain) {
            if (q_next >= q_end) {
                unsigned h = 0;
                if (lane == 0 && skip < n_in) h = atomicAdd(&a.counters[CNT_WORK + a.round], (unsigned)QW_GRAB);
                h = (unsigned)uniform((int)__shfl((int)h, 0));
                q_next = skip < n_in ? skip + h : n_in;
                q_end = q_next + QW_GRAB;
            }
            qi = q_next++;
        }
        if (qi >= n_in) break;
        float qx, qy, qz;
        int g, c = 0;
        if (plain) {
            qx = a.q_xyz[3 * (size_t)qi]; qy = a.q_xyz[3 * (size_t)qi + 1]; qz = a.q_xyz[3 * (size_t)qi + 2];
            g = (int)qi;
        } else {
            const int s = a.base + (int)qi;
            const float4 q = a.qs[s];
            qx = q.x; qy = q.y; qz = q.z;
            g = __float_as_int(q.w);
            c = a.qcloud ? a.qcloud[s] : 0;
        }
        g = uniform(g);
        c = uniform(c);
        const int k = uniform(a.kc ? a.kc[c] : a.k);
        const int2 root = a.roots[c];
        int round = a.round;
        float R = plain ? INFINITY : a.radius0[c];
        for (int r = 0; r < round; ++r) R *= a.mult;
        WaveQuery<M> Q;
        Q.qx = qx; Q.qy = qy; Q.qz = qz;
        Q.tests = 0;
        bool ok = true;
        while (true) {
            const bool last = a.last || round == a.rounds - 1;
            Q.T0 = last ? INFINITY : (double)(R * R);
            Q.L.init(k, lane);
            Q.wd = INFINITY;
            Q.wi = GPRT_ID_NONE;
            ok = qwave_walk<M>(Q, nodes, a.pts, make_int2(uniform(root.x), uniform(root.y)), lane, a.max_steps,
                               visits);
            if (!ok || last || Q.wd <= Q.T0) break;
            ++round;
            R *= a.mult;
            if (lane == 0) atomicAdd(&a.counters[CNT_ACTIVE + round], 1u);  // statistics
        }
        tests += Q.tests;
        if (!ok) {
            if (lane == 0) atomicAdd(&a.counters[CNT_ERROR], 1u);
        } else if (Q.wd == INFINITY && lane == 0) {
            atomicAdd(&a.counters[CNT_PARTIAL], 1u);
        }
        const int64_t ob = a.out_base ? a.out_base[c] + (int64_t)(g - a.qoff[c]) * k : (int64_t)g * k;
        const int s0 = 64 * M - k;
#pragma unroll
        for (int p = 0; p < M; ++p) {
            const int j = lane * M + p - s0;
            if (j >= 0) {
                a.ids_out[ob + j] = Q.L.id[p];
                if (a.d2_out) a.d2_out[ob + j] = Q.L.d[p];
            }
        }
        if (plain) qi += nwaves;
    }
    if (COUNT && lane == 0) {
        atomicAdd(&a.stats64[0], (unsigned long long)visits);
        atomicAdd(&a.stats64[1], (unsigned long long)tests);
    }
}

// list slots per lane for a largest k
int qwave_lanes_m(int kmax) {
    int m = 1;
    while (64 * m < kmax) m <<= 1;
    return m;
}

// The reference's register-heap program (gprt-macros/src/lib.rs:305-377) for
// k > 64: per query (input order, radius in q.w) the k nearest points with
// d64 < r^2 (strict, as its AnyHit insert), written as ids[0..k) then the
// distances [k..2k) as float bits; empty slots hold 0xFFFFFFFF and fl32(r*r);
// lens = number found.  Live slots start at (r^2, 0), so a point enters only
// below r^2.
template <int M>
__global__ __launch_bounds__(256) void heap_wave_kernel(const Node* __restrict__ nodes_, const float4* __restrict__ pts,
                                                        int2 root, const float4* __restrict__ q4, int nq, int k,
                                                        unsigned* __restrict__ heap, unsigned* __restrict__ lens,
                                                        int max_steps, unsigned* err) {
    const int lane = threadIdx.x & 63;
    const GPRT_AS4 Node* nodes = const_view(nodes_);
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    uint64_t visits = 0;
    for (int qi = uniform((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6)); qi < nq; qi += nwaves) {
        const float4 q = q4[qi];
        const double r2 = (double)q.w * (double)q.w;
        WaveQuery<M> Q;
        Q.qx = q.x; Q.qy = q.y; Q.qz = q.z;
        Q.T0 = INFINITY;
        Q.tests = 0;
        Q.L.init(k, lane, r2, 0u);
        Q.L.worst(Q.wd, Q.wi);
        if (!qwave_walk<M>(Q, nodes, pts, root, lane, max_steps, visits) && lane == 0) atomicAdd(err, 1u);
        const float rsq = q.w * q.w;
        unsigned* h = heap + (size_t)qi * 2 * k;
        const int s0 = 64 * M - k;
        unsigned cnt = 0;
#pragma unroll
        for (int p = 0; p < M; ++p) {
            const int o = lane * M + p - s0;
            const bool real = o >= 0 && Q.L.d[p] < r2;
            if (o >= 0) {
                h[o] = real ? Q.L.id[p] : 0xFFFFFFFFu;
                h[k + o] = real ? __float_as_uint((float)Q.L.d[p]) : __float_as_uint(rsq);
            }
            cnt += real;
        }
        for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
        if (lane == 0) lens[qi] = cnt;
    }
}

int launch_heap_wave(hipStream_t s, const Node* nodes, const float4* pts, int2 root, const float4* q4, int nq, int k,
                     unsigned* heap, unsigned* lens, int max_steps, unsigned* err) {
    if (k > GPRT_HIP_K_MAX) return fail(GPRT_ERR_ARG, "k too large for the wave-per-query kernel");
    const dim3 grid((nq + 3) / 4);  // one wavefront per query
    switch (qwave_lanes_m(k)) {
#define GPRT_HW(MM)                                                                                                     case MM:                                                                                                                hipLaunchKernelGGL(heap_wave_kernel<MM>, grid, dim3(256), 0, s, nodes, pts, root, q4, nq, k, heap, lens,                             max_steps, err);                                                                                 break;
        GPRT_HW(1) GPRT_HW(2) GPRT_HW(4) GPRT_HW(8) GPRT_HW(16) GPRT_HW(32)
        default: GPRT_HW(64)
#undef GPRT_HW
    }
    return hipGetLastError() == hipSuccess ? GPRT_OK : fail(GPRT_ERR_HIP, "heap program launch failed");
}

template <int M>
static int qwave_grid(int grid_req) {
    if (grid_req > 0) return grid_req;
    static std::atomic<int> cache[64];
    return resident_grid(knn_qwave_kernel<M, false>, cache);
}

int launch_qwave(hipStream_t s, const RoundArgs& a, int kmax, int grid_req, bool count) {
    if (kmax > GPRT_HIP_K_MAX) return fail(GPRT_ERR_ARG, "k too large for the wave-per-query kernel");
    const int m = qwave_lanes_m(kmax);
#define GPRT_QW(MM)                                                                                         \
    {                                                                                                       \
        const int grid = qwave_grid<MM>(grid_req);                                                          \
        if (count) hipLaunchKernelGGL((knn_qwave_kernel<MM, true>), dim3(grid), dim3(256), 0, s, a);       \
        else hipLaunchKernelGGL((knn_qwave_kernel<MM, false>), dim3(grid), dim3(256), 0, s, a);            \
    }
    switch (m) {
        case 1: GPRT_QW(1) break;
        case 2: GPRT_QW(2) break;
        case 4: GPRT_QW(4) break;
        case 8: GPRT_QW(8) break;
        case 16: GPRT_QW(16) break;
        case 32: GPRT_QW(32) break;
        default: GPRT_QW(64) break;
    }
#undef GPRT_QW
    return hipGetLastError() == hipSuccess ? GPRT_OK : fail(GPRT_ERR_HIP, "wave-per-query launch failed");
}

}  // namespace gprt
