// knn_packet.hip -- one TrueKNN round (gprt-macros/src/lib.rs:466-594) as a
// persistent launch walked by 64-query wavefront packets.
#include <algorithm>

#include "knn_device.hpp"

namespace gprt {

// ---------------------------------------------------------------------------
// one bounded (or the final unbounded) round as a persistent launch.  Each
// wavefront takes 64 consecutive queries of the active list (Morton order, so
// they are spatial neighbours) and walks the tree as a packet: node and leaf
// point loads are wave-uniform (scalar loads), every lane tests them against its
// own query and its own threshold, and a child is visited when any lane needs
// it.  The traversal stack is one uniform LDS stack per wavefront; the radix
// tree depth bound keeps it from overflowing.
// ---------------------------------------------------------------------------
constexpr int PACKET_STACK = EXACT_STACK;

template <int C, int G, bool COUNT>
__global__ __launch_bounds__(256) void knn_packet_kernel(RoundArgs a) {
    __shared__ int stk[4][PACKET_STACK];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned n_in = a.counters[CNT_ACTIVE + a.round];
    if (n_in == 0) return;
    const GPRT_AS4 Node* nodes = const_view(a.nodes);      // uniform node loads
    const GPRT_AS4 float4* upts = const_view(a.pts);       // uniform leaf-point loads
    const GPRT_AS4 int2* roots = const_view(a.roots);
    const float4* __restrict__ pts = a.pts;                // per-lane loads (re-rank)
    uint64_t visits = 0, tests = 0, reranks = 0;
    const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nwaves = (gridDim.x * blockDim.x) >> 6;
    bool first = true;
    while (true) {
        unsigned base = 0;
        if (lane == 0) base = grab_work_first(&a.counters[CNT_WORK + a.round], n_in, first, wave, nwaves);
        base = __shfl(base, 0);
        first = false;
        if (base >= n_in) break;
        const unsigned i = base + lane;
        const bool valid = i < n_in;
        const int s = valid ? (a.in_list ? a.in_list[i] : a.base + (int)i) : 0;
        const float4 q = valid ? a.qs[s] : make_float4(0.f, 0.f, 0.f, 0.f);
        const int c = valid && a.qcloud ? a.qcloud[s] : 0;
        const int k = a.kc ? a.kc[c] : a.k;
        float R = a.last ? INFINITY : a.radius0[c];
        if (!a.last)
            for (int r = 0; r < a.round; ++r) R *= a.mult;
        const float T0 = a.last ? INFINITY : R * R;
        F32Query<C, G> Q;
        Q.start(q.x, q.y, q.z, T0, k, false);
        bool overflow = false;
        // lanes of one wave may belong to different clouds (at cloud borders of
        // the sorted order): walk once per distinct cloud
        uint64_t todo = __ballot(valid);
        while (todo) {
            const int c0 = uniform(__shfl(c, __ffsll((unsigned long long)todo) - 1));
            const bool mine = valid && c == c0;
            todo &= ~__ballot(mine);
            if (mine) Q.tau = band(T0);
            const int2 root = ld2i(roots + c0);
            if (root.y > 0 && root.x < 0) {
                const int f = uniform(~root.x);
                for (int t = 0; t < root.y; ++t) Q.visit_point(ld4(upts + f + t), f + t);
                if (COUNT) tests += root.y;
            }
            if (root.y <= 0 || root.x < 0) {
                Q.tau = -1.0f;  // this cloud's lanes are done
                continue;
            }
            int sp = 0;
            int node = uniform(root.x);
            int steps = 0;
            while (true) {
                if (++steps > a.max_steps) { overflow = true; break; }
                const Node nd = ldnode(nodes + node);
                if (COUNT) ++visits;
                const float dA = box2f(Q.qx, Q.qy, Q.qz, nd.alo, nd.ahi);
                const float dB = box2f(Q.qx, Q.qy, Q.qz, nd.blo, nd.bhi);
                const int rA = uniform(__float_as_int(nd.alo.w)), rB = uniform(__float_as_int(nd.blo.w));
                const bool needA = dA <= Q.tau, needB = dB <= Q.tau;
                const bool anyA = __ballot(needA) != 0, anyB = __ballot(needB) != 0;
                // lanes vote for the child nearer to them
                const int votesB = __popcll(__ballot(needB && (!needA || dB < dA)));
                const int votesA = __popcll(__ballot(needA && (!needB || dA <= dB)));
                const bool bfirst = votesB > votesA;
                // leaf buckets are scanned at once
                if (rA < 0 && rB < 0 && bfirst) {
                    if (anyB) {
                        const int f = ~rB, m = uniform(__float_as_int(nd.bhi.w));
                        for (int t = 0; t < m; ++t) Q.visit_point(ld4(upts + f + t), f + t);
                        if (COUNT) tests += m;
                    }
                    if (anyA) {
                        const int f = ~rA, m = uniform(__float_as_int(nd.ahi.w));
                        for (int t = 0; t < m; ++t) Q.visit_point(ld4(upts + f + t), f + t);
                        if (COUNT) tests += m;
                    }
                } else {
                    if (rA < 0 && anyA) {
                        const int f = ~rA, m = uniform(__float_as_int(nd.ahi.w));
                        for (int t = 0; t < m; ++t) Q.visit_point(ld4(upts + f + t), f + t);
                        if (COUNT) tests += m;
                    }
                    if (rB < 0 && anyB) {
                        const int f = ~rB, m = uniform(__float_as_int(nd.bhi.w));
                        for (int t = 0; t < m; ++t) Q.visit_point(ld4(upts + f + t), f + t);
                        if (COUNT) tests += m;
                    }
                }
                const bool goA = rA >= 0 && anyA, goB = rB >= 0 && anyB;
                if (goA && goB) {
                    if (sp >= PACKET_STACK) { overflow = true; break; }
                    if (lane == 0) stk[w][sp] = bfirst ? rA : rB;
                    ++sp;
                    node = bfirst ? rB : rA;
                    continue;
                }
                if (goA) { node = rA; continue; }
                if (goB) { node = rB; continue; }
                if (sp == 0) break;
                --sp;
                __builtin_amdgcn_wave_barrier();
                node = uniform(stk[w][sp]);
            }
            Q.tau = -1.0f;  // this cloud's lanes are done
            if (overflow) break;
        }
        // resolve
        int status = 0;  // 0 done, 1 unresolved, 2 fallback
        float fbound = INFINITY;  // fallback: an upper bound on the k-th d64
        if (valid) {
            const float kth = Q.kth();
            const bool full = kth < INFINITY;
            const bool resolved = a.last ? true : (full && kth <= T0);
            if (!resolved) {
                status = 1;
            } else {
                const float tf = band(kth);  // INF when the cloud holds fewer than k points
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
        const int slot = wave_append(&a.counters[CNT_ACTIVE + a.round + 1], status == 1);
        if (slot >= 0) a.out_list[slot] = s;
        const int fslot = wave_append(&a.counters[CNT_FALLBACK], status == 2);
        if (fslot >= 0) a.fb_list[fslot] = make_int2(s, __float_as_int(fbound));
    }
    if (COUNT) {
        // visits/tests are wave-uniform in the packet walk: count once per wave
        for (int o = 32; o > 0; o >>= 1) reranks += __shfl_xor(reranks, o);
        if (lane == 0) {
            atomicAdd(&a.stats64[0], (unsigned long long)visits);
            atomicAdd(&a.stats64[1], (unsigned long long)tests);
            atomicAdd(&a.stats64[2], (unsigned long long)reranks);
        }
    }
}


void launch_packet_round(hipStream_t s, const RoundArgs& a, int cls, int grid, bool count) {
    if (grid <= 0) grid = 1024;
#define GPRT_LP(CAP, G)                                                                              \
    if (count) hipLaunchKernelGGL((knn_packet_kernel<CAP, G, true>), dim3(grid), dim3(256), 0, s, a); \
    else hipLaunchKernelGGL((knn_packet_kernel<CAP, G, false>), dim3(grid), dim3(256), 0, s, a);
    switch (cls) {
        case 0: GPRT_LP(8, GUARD) break;
        case 1: GPRT_LP(16, GUARD) break;
        case 2: GPRT_LP(24, GUARD) break;
        case 3: GPRT_LP(32, GUARD) break;
        case 4: GPRT_LP(40, GUARD) break;
        default: GPRT_LP(48, GUARD) break;  // class 5
    }
#undef GPRT_LP
}

}  // namespace gprt
