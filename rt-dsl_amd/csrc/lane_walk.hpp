// lane_walk.hpp -- the per-lane FP32 k-NN walk with its fixed-slot candidate
// list and FP64 re-rank, as __host__ __device__ code.
//
// knn_lane_kernel (knn_lane.hip) runs it on the GPU with the LDS ring and the
// HBM spill column of its lane; the host build of the same code
// (host_walk.cpp) runs it one query at a time with bounds-checked memory, so
// the walk's indexing is checked on the CPU.  The memory policy `Mem` supplies
//   turn(scanning)                                   -- whether the lane takes this trip
//   fetch4(scanning, first_point, node, last, v[4])  -- the trip's 64-byte load
//   ring_ref(j) / ring_d(j) / ring_put(j, ref, d)     -- LDS ring slot j
//   spill_get(level) / spill_put(level, int2)         -- HBM spill of the lane
#pragma once
#include "gprt_internal.hpp"

namespace gprt {

// spare candidate slots beyond k (the guard band)
constexpr int GUARD = 3;

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
    GPRT_HD void init(int sentinels) {
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
    GPRT_HD float insert(float dd, int ss) {
        const float ev = d[C - 1];
#pragma unroll
        for (int j = C - 1; j >= 1; --j) {
            const bool shift = d[j - 1] > dd;
            const bool here = d[j] > dd;
            s[j] = shift ? s[j - 1] : (here ? ss : s[j]);
            d[j] = med3f(d[j - 1], d[j], dd);
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

    GPRT_HD void start(float x, float y, float z, float t0, int k, bool active) {
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
    GPRT_HD void visit_point(float4 p, int s) {
        const float dd = d2f(qx, qy, qz, p.x, p.y, p.z);
        if (dd <= tau) {
            const bool ins = dd < L.d[C - 1];
            const float ev = L.insert(ins ? dd : INFINITY, s);
            ev_min = fminf(ev_min, ins ? ev : dd);
            T = fminf(T0, L.d[C - 1 - G]);
            tau = band(T);
        }
    }
    GPRT_HD float kth() const { return L.d[C - 1 - G]; }
};

// ---------------------------------------------------------------------------
// One walk of one query (closest child first), in trips of exactly one
// 64-byte load: either the next node (two child boxes) or the next four points
// of the leaf range being scanned.  A node's in-range leaf buckets (the two
// buckets of a node are adjacent ranges of the sorted points, so both form one
// range) are scanned over the following trips, and only then is the node's
// successor chosen, with the threshold those points left.  The stack holds
// (ref, d32 of the box) pairs: the S most recent in the ring, older ones in the
// spill column; a popped entry already beyond the threshold is dropped without
// a node load.  Returns the number of trips; `overflow` is set when the walk
// outgrew max_depth stack entries or max_steps node visits.
// ---------------------------------------------------------------------------
template <int C, int G, int S, bool COUNT, class Mem>
GPRT_HD int lane_walk(F32Query<C, G>& Q, int2 root, const Mem& m, int max_steps, int max_depth, bool& overflow,
                      uint64_t& visits, uint64_t& tests) {
    int sp = 0, lo = 0;  // stack depth; entries [0, lo) are in the spill column
    int steps = 0, trips = 0;
    // walk state: the node to load next, or a leaf range [sf, sf + sm)
    // whose points from sf + st on are still to be scanned
    bool scanning = root.x < 0;
    int node = root.x;
    int sf = scanning ? ~root.x : 0, sm = scanning ? root.y : 0, st = 0;
    // children of the last node: box distances and refs (negative
    // refs are leaf buckets, never descended into)
    float cdA = INFINITY, cdB = INFINITY;
    int crA = -1, crB = -1;
    while (true) {
        // m.turn: whether this lane takes the trip (the device build may let
        // lanes of one kind, node or scan, wait for the others; each lane's
        // own sequence of trips is the same either way)
        if (!m.turn(scanning)) continue;
        ++trips;
        const int avail = scanning ? sm - st : 4;
        float4 v[4];
        // four loads without branches (so they issue back to back whichever
        // lanes scan): past the end of a leaf range the index is clamped to
        // its last point, which is loaded but not visited
        m.fetch4(scanning, sf + st, node, scanning ? avail - 1 : 3, v);
        if (!scanning) {
            if (++steps > max_steps) { overflow = true; break; }
            if (COUNT) ++visits;
            // v = child A's box (lo, hi), child B's box (lo, hi)
            cdA = box2f(Q.qx, Q.qy, Q.qz, v[0], v[1]);
            cdB = box2f(Q.qx, Q.qy, Q.qz, v[2], v[3]);
            crA = as_int(v[0].w);
            crB = as_int(v[2].w);
            const bool sA = crA < 0 && cdA <= Q.tau, sB = crB < 0 && cdB <= Q.tau;
            sf = sA ? ~crA : ~crB;
            sm = (sA ? as_int(v[1].w) : 0) + (sB ? as_int(v[3].w) : 0);
            st = 0;
            if (sm > 0) {
                if (COUNT) tests += sm;
                scanning = true;
                continue;
            }
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (u < avail) Q.visit_point(v[u], sf + st + u);
            st += 4;
            if (st < sm) continue;
            scanning = false;
        }
        // the last node's successor, with the threshold its buckets left
        const bool gA = crA >= 0 && cdA <= Q.tau, gB = crB >= 0 && cdB <= Q.tau;
        const bool bfirst = cdB < cdA;
        if (gA && gB) {
            if (sp >= max_depth) { overflow = true; break; }
            if (sp - lo == S) {  // ring full: its oldest entry goes to the spill column
                const int j = lo & (S - 1);
                m.spill_put(lo, make_int2(m.ring_ref(j), as_int(m.ring_d(j))));
                ++lo;
            }
            m.ring_put(sp & (S - 1), bfirst ? crA : crB, bfirst ? cdA : cdB);
            ++sp;
            node = bfirst ? crB : crA;
            continue;
        }
        if (gA) { node = crA; continue; }
        if (gB) { node = crB; continue; }
        // pop the nearest-pushed entry still inside the threshold
        node = -1;
        while (sp > 0) {
            --sp;
            int r;
            float d;
            if (sp >= lo) {
                r = m.ring_ref(sp & (S - 1));
                d = m.ring_d(sp & (S - 1));
            } else {
                const int2 e = m.spill_get(sp);
                lo = sp;
                r = e.x;
                d = as_float(e.y);
            }
            if (d <= Q.tau) { node = r; break; }
        }
        if (node < 0) break;
    }
    return trips;
}

// FP64 re-rank of the guard band: candidates are the list entries with
// 0 <= d32 <= tf (sentinels are -inf); the result is (dd, id) sorted by
// (d64, id), the non-candidates as (+inf, GPRT_ID_NONE) at the end.  `pt(i)`
// loads sorted point i.  Every candidate's sorted position must lie in its
// cloud's range [lo, hi) and its id below hi - lo: `bad` reports a violation
// (a candidate outside the range is read at lo instead).  Returns the number
// of candidates.
template <int C, class PtFn>
GPRT_HD int rerank_sort(const float* ld, const int* ls, float qx, float qy, float qz, float tf, int64_t lo,
                        int64_t hi, const PtFn& pt, double* dd, uint32_t* id, bool& bad) {
    int reads = 0;
#pragma unroll
    for (int j = 0; j < C; ++j) {
        const bool in = ld[j] >= 0.0f && ld[j] <= tf && ls[j] >= 0;
        dd[j] = INFINITY;
        id[j] = GPRT_ID_NONE;
        reads += in;
        if (in) {
            const bool out = ls[j] < lo || ls[j] >= hi;
            const float4 p = pt(out ? lo : (int64_t)ls[j]);
            dd[j] = d2d(qx, qy, qz, p.x, p.y, p.z);
            id[j] = (uint32_t)as_int(p.w);
            bad |= out || (int64_t)id[j] >= hi - lo;
        }
    }
    bool swapped = true;
    while (swapped) {
        swapped = false;
#pragma unroll
        for (int j = 0; j + 1 < C; ++j) {
            const bool gt = dd[j] > dd[j + 1] || (dd[j] == dd[j + 1] && id[j] > id[j + 1]);
            const double a = gt ? dd[j + 1] : dd[j], b = gt ? dd[j] : dd[j + 1];
            const uint32_t ia = gt ? id[j + 1] : id[j], ib = gt ? id[j] : id[j + 1];
            dd[j] = a; dd[j + 1] = b; id[j] = ia; id[j + 1] = ib;
            swapped |= gt;
        }
    }
    return reads;
}

}  // namespace gprt
