// knn_wide.hip -- one TrueKNN round (gprt-macros/src/lib.rs:466-594) as a
// persistent launch with one query per lane over the 4-wide view of the LBVH.
#include <algorithm>
#include <atomic>

#include "knn_device.hpp"

namespace gprt {

// minimum wavefronts per SIMD the register allocation of the capacity-8
// class must allow (1: no constraint)
#ifndef GPRT_WIDE_MIN_WAVES
#define GPRT_WIDE_MIN_WAVES 1
#endif

// ---------------------------------------------------------------------------
// The per-lane round over the 4-wide view: one node visit tests four child
// boxes (half the depth of the binary walk, half the dependent node loads).
// In-range leaf slots are scanned at once; consecutive slots cover adjacent
// point ranges, so runs of in-range leaf slots are single ranges, at most two
// per node, scanned from one code site.  In-range internal slots are ordered
// by box distance (a 5-comparator network): the walk descends into the
// nearest and pushes the others farthest first onto the (ref, d32) LDS ring
// with HBM spill of knn_lane_kernel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cswap(float& da, int& ra, float& db, int& rb) {
    const bool sw = db < da;
    const float td = sw ? db : da, tu = sw ? da : db;
    const int rd = sw ? rb : ra, ru = sw ? ra : rb;
    da = td; db = tu; ra = rd; rb = ru;
}

template <int C, int G, int S, bool COUNT, bool ROUND0>
__global__ __launch_bounds__(256, C == 8 ? GPRT_WIDE_MIN_WAVES : 1) void knn_wide_kernel(RoundArgs a) {
    static_assert((S & (S - 1)) == 0, "LDS ring size must be a power of two");
    __shared__ int sref[S][256];
    __shared__ float sdist[S][256];
    const int tid = threadIdx.x, lane = tid & 63;
    const unsigned n_in = a.counters[CNT_ACTIVE + a.round];
    if (n_in == 0) return;
    const Node4* __restrict__ nodes = a.nodes4;
    const float4* __restrict__ pts = a.pts;
    const size_t gstride = (size_t)gridDim.x * 256;
    int2* __restrict__ spill = a.spill + (size_t)blockIdx.x * 256 + tid;
    const int max_depth = S + a.spill_depth;
    uint64_t visits = 0, tests = 0, reranks = 0, lane_steps = 0, wave_steps = 0;
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
            if (root.y > 0 && root.x < 0) {
                scan_bucket(Q, pts, ~root.x, root.y);
                if (COUNT) tests += root.y;
            } else if (root.y > 0) {
                int sp = 0, lo = 0;  // stack depth; entries [0, lo) are in HBM
                int node = root.x, steps = 0;
                while (true) {
                    if (++steps > a.max_steps) { overflow = true; break; }
                    const Node4 nd = nodes[node];
                    if (COUNT) ++visits;
                    float d0 = d2f(Q.qx, Q.qy, Q.qz, clampf(Q.qx, nd.lox.x, nd.hix.x), clampf(Q.qy, nd.loy.x, nd.hiy.x),
                                   clampf(Q.qz, nd.loz.x, nd.hiz.x));
                    float d1 = d2f(Q.qx, Q.qy, Q.qz, clampf(Q.qx, nd.lox.y, nd.hix.y), clampf(Q.qy, nd.loy.y, nd.hiy.y),
                                   clampf(Q.qz, nd.loz.y, nd.hiz.y));
                    float d2 = d2f(Q.qx, Q.qy, Q.qz, clampf(Q.qx, nd.lox.z, nd.hix.z), clampf(Q.qy, nd.loy.z, nd.hiy.z),
                                   clampf(Q.qz, nd.loz.z, nd.hiz.z));
                    float d3 = d2f(Q.qx, Q.qy, Q.qz, clampf(Q.qx, nd.lox.w, nd.hix.w), clampf(Q.qy, nd.loy.w, nd.hiy.w),
                                   clampf(Q.qz, nd.loz.w, nd.hiz.w));
                    if (nd.cnt.y == 0) d1 = INFINITY;
                    if (nd.cnt.w == 0) d3 = INFINITY;
                    // runs of in-range leaf slots: at most two contiguous point ranges
                    {
                        const float tau = Q.tau;
                        const bool l0 = nd.ref.x < 0 && d0 <= tau, l1 = nd.ref.y < 0 && d1 <= tau;
                        const bool l2 = nd.ref.z < 0 && d2 <= tau, l3 = nd.ref.w < 0 && d3 <= tau;
                        int f0 = 0, m0 = 0, f1 = 0, m1 = 0;
                        // first run starts at the first in-range leaf slot; the second
                        // at the first in-range leaf slot after a gap
                        bool in_run = false, second = false;
#define GPRT_RUN(L, R, N)                                      \
                        if (L) {                               \
                            if (!second && (in_run || m0 == 0)) { \
                                if (m0 == 0) f0 = ~(R);        \
                                m0 += (N);                     \
                            } else {                           \
                                if (m1 == 0) f1 = ~(R);        \
                                m1 += (N);                     \
                                second = true;                 \
                            }                                  \
                            in_run = true;                     \
                        } else {                               \
                            if (in_run && m0 > 0) second = true; \
                            in_run = false;                    \
                        }
                        GPRT_RUN(l0, nd.ref.x, nd.cnt.x)
                        GPRT_RUN(l1, nd.ref.y, nd.cnt.y)
                        GPRT_RUN(l2, nd.ref.z, nd.cnt.z)
                        GPRT_RUN(l3, nd.ref.w, nd.cnt.w)
#undef GPRT_RUN
#pragma nounroll
                        for (int run = 0; run < 2; ++run) {
                            const int f = run ? f1 : f0, m = run ? m1 : m0;
                            if (m > 0) {
                                scan_bucket(Q, pts, f, m);
                                if (COUNT) tests += m;
                            }
                        }
                    }
                    // internal slots still in range, nearest first
                    const float tau = Q.tau;
                    float e0 = (nd.ref.x >= 0 && d0 <= tau) ? d0 : INFINITY;
                    float e1 = (nd.ref.y >= 0 && d1 <= tau) ? d1 : INFINITY;
                    float e2 = (nd.ref.z >= 0 && d2 <= tau) ? d2 : INFINITY;
                    float e3 = (nd.ref.w >= 0 && d3 <= tau) ? d3 : INFINITY;
                    int r0 = nd.ref.x, r1 = nd.ref.y, r2 = nd.ref.z, r3 = nd.ref.w;
                    cswap(e0, r0, e1, r1);
                    cswap(e2, r2, e3, r3);
                    cswap(e0, r0, e2, r2);
                    cswap(e1, r1, e3, r3);
                    cswap(e1, r1, e2, r2);
                    if (e0 < INFINITY) {
                        // push the in-range others, farthest first
                        const int np = (e1 < INFINITY) + (e2 < INFINITY) + (e3 < INFINITY);
                        if (sp + np > max_depth) { overflow = true; break; }
#pragma nounroll
                        for (int j = 3; j >= 1; --j) {
                            const float ej = j == 3 ? e3 : (j == 2 ? e2 : e1);
                            const int rj = j == 3 ? r3 : (j == 2 ? r2 : r1);
                            if (ej < INFINITY) {
                                if (sp - lo == S) {  // ring full: its oldest entry goes to HBM
                                    const int jj = lo & (S - 1);
                                    spill[lo * gstride] = make_int2(sref[jj][tid], __float_as_int(sdist[jj][tid]));
                                    ++lo;
                                }
                                sref[sp & (S - 1)][tid] = rj;
                                sdist[sp & (S - 1)][tid] = ej;
                                ++sp;
                            }
                        }
                        node = r0;
                        continue;
                    }
                    // pop the nearest-pushed entry still inside the threshold
                    node = -1;
                    while (sp > 0) {
                        --sp;
                        int r;
                        float d;
                        if (sp >= lo) {
                            r = sref[sp & (S - 1)][tid];
                            d = sdist[sp & (S - 1)][tid];
                        } else {
                            const int2 e = spill[sp * gstride];
                            lo = sp;
                            r = e.x;
                            d = __int_as_float(e.y);
                        }
                        if (d <= Q.tau) { node = r; break; }
                    }
                    if (node < 0) break;
                }
                task_steps += steps;
            }
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
    return resident_grid(knn_wide_kernel<C, G, S, false, true>, cache);
}

int wide_grid(int cls, int ring, int grid_req) {
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
        if (count) hipLaunchKernelGGL((knn_wide_kernel<C, G, S, true, true>), dim3(grid), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((knn_wide_kernel<C, G, S, false, true>), dim3(grid), dim3(256), 0, s, a);
    } else {
        if (count) hipLaunchKernelGGL((knn_wide_kernel<C, G, S, true, false>), dim3(grid), dim3(256), 0, s, a);
        else hipLaunchKernelGGL((knn_wide_kernel<C, G, S, false, false>), dim3(grid), dim3(256), 0, s, a);
    }
}

void launch_wide_round(hipStream_t s, const RoundArgs& a, int cls, int ring, int grid_req, bool count) {
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
