// range.hip -- fixed-radius range query over the LBVH (r_nn! / range_query!,
// gprt-macros/src/lib.rs:729-750; the predicate is Sphere::intersect's
// d^2 <= r^2, gprt-core/src/lib.rs:82-93).
//
// Contract: query q returns every point with d64(q,p) <= (double)r*(double)r,
// ordered by (d64, id), as a CSR list.  The reference's version returns one
// flat array of at most 1e6 ids for all queries with no per-query split
// (lib.rs:662-668); this one keeps the split.
//
// Pipeline: Morton-ordered queries -> FP32 traversal entering boxes within
// the guard threshold band(fl32(r*r)), points confirmed in FP64, subtrees
// proved inside the ball (inner_band) taken whole without a visit -> per-query
// counts -> offsets (exclusive scan) -> the same traversal writing slots
// (point positions, and (first, count) runs for whole subtrees) -> per-query
// expansion of the runs, ids and d64 from the gathered points, sort by
// (d64, id).
#include <algorithm>
#include <atomic>
#include <cfloat>

#include "gprt_internal.hpp"
#include "knn_internal.hpp"

namespace gprt {

// Output slots of the fill walk.  A query's range [off[q], off[q+1]) holds one
// entry per hit in the count, but the walk writes fewer slots: a hit found by
// a point test is one slot holding its sorted point position; a subtree whose
// box lies wholly inside the ball (every point of it a hit, see inner_band) is
// two slots, RUN_START | first position and RUN_LEN | number of points.
// Positions and lengths are below 2^30 (the radix sort's key limit).  The
// number of slots written is kept per query (`used`); the sort kernels expand
// the runs and overwrite the range with ids.
constexpr uint32_t RUN_START = 0x80000000u, RUN_LEN = 0x40000000u, SLOT_MASK = 0x3FFFFFFFu;

// elements a slot stands for: a position 1, a run start the length held by the
// next slot, a length slot 0
__device__ __forceinline__ uint32_t slot_elems(uint32_t v, uint32_t next) {
    return (v & RUN_START) ? (next & SLOT_MASK) : ((v & RUN_LEN) ? 0u : 1u);
}

// query g's output range [b, e), clamped into [0, capacity): offsets that do
// not match the walk never move a write out of the output array
__device__ __forceinline__ void seg_bounds(const int64_t* __restrict__ off, int g, int64_t capacity, int64_t& b,
                                           int64_t& e) {
    b = min(max(off[g], (int64_t)0), capacity);
    e = min(max(off[g + 1], b), capacity);
}

// Per-lane walk with a private stack of EXACT_STACK entries (never overflows:
// the radix-tree depth bound).  A child box within band(r^2) is entered; one
// whose farthest corner is within inner_band(r^2) is taken whole: its points
// are the sorted positions [first, first + count), first being ~ref for a leaf
// bucket, ref - count + 1 for an internal A child and ref for an internal B
// child (a Karras node i covers [lo, hi] with its A child ending at split
// gamma = ref A and its B child starting at gamma + 1 = ref B).  The count pass
// stores each query's number of hits; the fill pass writes slots (above) in
// walk order and checks that the hits fill the range exactly (a range that
// does not match the queries and radius is counted in `bad`).  COUNT adds
// node visits, point tests and slots to stats64.
constexpr int RANGE_RING = 16;  // stack entries per lane kept in LDS by the range walk

template <bool WRITE, bool COUNT>
__global__ __launch_bounds__(256) void range_kernel(const Node* __restrict__ nodes, const float4* __restrict__ pts,
                                                    int2 root, const float4* __restrict__ qs, int nq, float r_all,
                                                    const float* __restrict__ radii,
                                                    int64_t* __restrict__ counts, const int64_t* __restrict__ off,
                                                    uint32_t* __restrict__ out, int64_t capacity,
                                                    uint32_t* __restrict__ used, int max_steps, unsigned* err,
                                                    unsigned* bad,
                                                    unsigned long long* stats64) {
    __shared__ int ring[RANGE_RING][256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t visits = 0, tests = 0, slots = 0;
    if (i < nq) {
        const float4 q = qs[i];
        const int g = __float_as_int(q.w);
        const float r = radii ? radii[g] : r_all;
        const double r2 = (double)r * (double)r;
        const float s2 = r * r;
        const float tau = band(s2), tin = inner_band(s2);
        int64_t hits = 0, beg = 0, end = 0;
        if (WRITE) seg_bounds(off, g, capacity, beg, end);
        int64_t pos = beg;
        auto single = [&](int p) {
            if (WRITE) {
                if (pos < end) out[pos] = (uint32_t)p;
                ++pos;
            }
            ++hits;
            if (COUNT) ++slots;
        };
        auto run = [&](int f, int c) {
            if (c == 1) {
                single(f);
                return;
            }
            if (WRITE) {
                if (pos + 1 < end) {
                    out[pos] = RUN_START | (uint32_t)f;
                    out[pos + 1] = RUN_LEN | (uint32_t)c;
                }
                pos += 2;
            }
            hits += c;
            if (COUNT) slots += 2;
        };
        // a leaf range four points at a time: the loads of a group issue
        // together (indices past the range clamped to its last point, loaded
        // but not tested)
        auto visit = [&](int f, int m) {
            if (COUNT) tests += (uint64_t)m;
            for (int t0 = 0; t0 < m; t0 += 4) {
                float4 p[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) p[u] = pts[f + min(t0 + u, m - 1)];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (t0 + u < m && d2f(q.x, q.y, q.z, p[u].x, p[u].y, p[u].z) <= tau &&
                        d2d(q.x, q.y, q.z, p[u].x, p[u].y, p[u].z) <= r2)
                        single(f + t0 + u);
            }
        };
        if (root.y > 0) {
            if (root.x < 0) {
                visit(~root.x, root.y);
            } else {
                // the stack: its RANGE_RING most recent entries in the lane's LDS
                // column, older ones in the private array (scratch memory)
                int stk[EXACT_STACK];
                int sp = 0, lo = 0, node = root.x, steps = 0;  // entries [lo, sp) are in the ring
                const int tid = threadIdx.x;
                while (true) {
                    if (++steps > max_steps || sp >= EXACT_STACK) { atomicAdd(err, 1u); break; }
                    const Node nd = nodes[node];
                    if (COUNT) ++visits;
                    const float dA = box2f(q.x, q.y, q.z, nd.alo, nd.ahi);
                    const float dB = box2f(q.x, q.y, q.z, nd.blo, nd.bhi);
                    const int rA = __float_as_int(nd.alo.w), rB = __float_as_int(nd.blo.w);
                    const int nA = __float_as_int(nd.ahi.w), nB = __float_as_int(nd.bhi.w);
                    const bool inA = dA <= tau, inB = dB <= tau;
                    const bool wA = inA && boxmax2f(q.x, q.y, q.z, nd.alo, nd.ahi) <= tin;
                    const bool wB = inB && boxmax2f(q.x, q.y, q.z, nd.blo, nd.bhi) <= tin;
                    if (wA || wB) {
                        // A's points are followed by B's: both whole are one run
                        const int fA = rA < 0 ? ~rA : rA - nA + 1;
                        const int fB = rB < 0 ? ~rB : rB;
                        run(wA ? fA : fB, (wA ? nA : 0) + (wB ? nB : 0));
                    }
                    {
                        // the two buckets of a node are adjacent ranges of the sorted
                        // points (A's then B's): in-range ones are scanned as one range
                        const bool sA = rA < 0 && inA && !wA, sB = rB < 0 && inB && !wB;
                        const int mm = (sA ? nA : 0) + (sB ? nB : 0);
                        if (mm > 0) visit(sA ? ~rA : ~rB, mm);
                    }
                    const bool gA = rA >= 0 && inA && !wA, gB = rB >= 0 && inB && !wB;
                    if (gA && gB) {
                        if (sp - lo == RANGE_RING) {  // ring full: its oldest entry to the array
                            stk[lo] = ring[lo & (RANGE_RING - 1)][tid];
                            ++lo;
                        }
                        ring[sp & (RANGE_RING - 1)][tid] = rB;
                        ++sp;
                        node = rA;
                        continue;
                    }
                    if (gA) { node = rA; continue; }
                    if (gB) { node = rB; continue; }
                    if (sp == 0) break;
                    --sp;
                    if (sp >= lo) {
                        node = ring[sp & (RANGE_RING - 1)][tid];
                    } else {
                        node = stk[sp];
                        lo = sp;
                    }
                }
            }
        }
        if (WRITE) {
            used[g] = (uint32_t)min(pos - beg, end - beg);
            if (hits != end - beg || off[g] != beg || off[g + 1] != end) atomicAdd(bad, 1u);
        } else {
            counts[g] = hits;
        }
    }
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) {
            visits += __shfl_xor(visits, o);
            tests += __shfl_xor(tests, o);
            slots += __shfl_xor(slots, o);
        }
        if ((threadIdx.x & 63) == 0) {
            atomicAdd(&stats64[0], (unsigned long long)visits);
            atomicAdd(&stats64[1], (unsigned long long)tests);
            atomicAdd(&stats64[2], (unsigned long long)slots);
        }
    }
}

// ---------------------------------------------------------------------------
// exclusive scan of int64 counts -> offsets[0..n] (three phases)
// ---------------------------------------------------------------------------
constexpr int SCAN_TILE = 1024;

__global__ __launch_bounds__(256) void scan_tile_sums(const int64_t* __restrict__ in, int n, int64_t* sums) {
    __shared__ int64_t red[256];
    int64_t s = 0;
    const int base = blockIdx.x * SCAN_TILE;
    for (int j = threadIdx.x; j < SCAN_TILE; j += 256)
        if (base + j < n) s += in[base + j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(1024) void scan_sums_kernel(int64_t* sums, int m, int64_t* total) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x, per = (m + 1023) / 1024;
    const int b = min(t * per, m), e = min(b + per, m);
    int64_t s = 0;
    for (int j = b; j < e; ++j) s += sums[j];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int64_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = t > 0 ? part[t - 1] : 0;
    for (int j = b; j < e; ++j) {
        int64_t v = sums[j];
        sums[j] = run;
        run += v;
    }
    if (t == 1023) *total = part[1023];
}

__global__ __launch_bounds__(256) void scan_tile_apply(const int64_t* __restrict__ in, int n,
                                                       const int64_t* __restrict__ sums, int64_t* __restrict__ out) {
    __shared__ int64_t v[SCAN_TILE];
    const int base = blockIdx.x * SCAN_TILE;
    for (int j = threadIdx.x; j < SCAN_TILE; j += 256) v[j] = base + j < n ? in[base + j] : 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t run = sums[blockIdx.x];
        for (int j = 0; j < SCAN_TILE; ++j) {
            int64_t x = v[j];
            v[j] = run;
            run += x;
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < SCAN_TILE; j += 256)
        if (base + j < n) out[base + j] = v[j];
}

__global__ void set_last_kernel(int64_t* out, int n, const int64_t* total) { out[n] = *total; }


// ---------------------------------------------------------------------------
// Per-query sort: the fill pass left each query's hits as sorted point
// positions in walk order; these kernels turn them into ids in (d64, id)
// order, in place.  d64 and the id are recomputed from the gathered point (a
// query's hits are a few runs of neighbouring positions, so the gathers are
// mostly cache hits) instead of being written and read back by the walk.
// A classifying pass lists the segments by length (wave-aggregated appends):
//   * up to 64 hits: one wavefront, a bitonic network in registers;
//   * up to SEG_MID (2,048) / SEG_LARGE (8,192): one workgroup of 256 / 1,024
//     lanes, a bucket sort.  With B = ceil(m/2) buckets, bucket b holds the
//     elements with floor(d64 * B / r^2) = b, a monotone function of d64, so
//     buckets come out in d64 order and only each bucket's few elements (two
//     on average) need ordering by (d64, id).  Each lane keeps its elements'
//     (d64, id) in registers, scatters them into bucket order in LDS, counts
//     the smaller keys of its bucket (keys are distinct: ids differ) and
//     writes each id straight to its output slot; a bucket of more than
//     SEG_RUN elements (many equal distances: lattices, duplicates) switches
//     the segment to a bitonic network over the whole segment;
//   * longer: SEG_LARGE-element chunks sorted the same way, then merged
//     pairwise in global memory by rank.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool pair_gt(double da, uint32_t ia, double db, uint32_t ib) {
    return da > db || (da == db && ia > ib);
}

constexpr int SEG_MID = 2048;
constexpr int SEG_LARGE = 8192;
constexpr int SEG_PARTS = 12288;  // longest segment sorted in LDS parts (longer: chunks merged in global memory)
constexpr int SEG_RUN = 32;       // longest bucket ordered by rank counting

struct SegQ {
    float x, y, z;
    double r2;
};
__device__ __forceinline__ SegQ seg_query(const float* __restrict__ qxyz, const float* __restrict__ radii,
                                          float r_all, int g) {
    const float r = radii ? radii[g] : r_all;
    return SegQ{qxyz[3 * (size_t)g], qxyz[3 * (size_t)g + 1], qxyz[3 * (size_t)g + 2], (double)r * (double)r};
}

// list indices of the per-length segment lists
enum { SEG_L_SMALL = 0, SEG_L_MID, SEG_L_LARGE, SEG_L_HUGE };

// one lane per segment: lists segments of 2..64, ..SEG_MID, ..SEG_LARGE and
// longer hits (one-hit segments are converted here); huge ones also get a range
// of the merge scratch (from the running total huge_total).  The list slots
// and scratch ranges are claimed once per workgroup and list: atomics from
// every wavefront on the four adjacent list counters serialise in one L2
// channel (about 20,000 of them took 0.29 ms for 400,000 segments).
constexpr int CLS_T = 1024;
__global__ __launch_bounds__(CLS_T) void seg_classify_kernel(const int64_t* __restrict__ off, int nq, int64_t capacity,
                                                             const float4* __restrict__ pts, uint32_t npts,
                                                             uint32_t* ids, int* lists, unsigned* counts,
                                                             int64_t* huge_tmp, unsigned long long* huge_total) {
    constexpr int W = CLS_T / 64;
    __shared__ unsigned wcnt[4][W];
    __shared__ unsigned long long whuge[W];
    __shared__ unsigned bbase[4];
    __shared__ unsigned long long hbase;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t m = 0, b = 0;
    if (g < nq) {
        int64_t e;
        seg_bounds(off, g, capacity, b, e);
        m = e - b;
        // a one-hit range holds one position (a run has two or more points)
        if (m == 1) ids[b] = (uint32_t)__float_as_int(pts[min(ids[b] & SLOT_MASK, npts - 1)].w);
    }
    const int cls = m <= 1 ? -1 : (m <= 64 ? SEG_L_SMALL : (m <= SEG_MID ? SEG_L_MID : (m <= SEG_LARGE ? SEG_L_LARGE
                                                                                                   : SEG_L_HUGE)));
    uint64_t mk[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        mk[c] = __ballot(cls == c);
        if (lane == 0) wcnt[c][w] = (unsigned)__popcll(mk[c]);
    }
    // the long segments' lengths: inclusive wavefront scan
    const uint64_t own = cls == SEG_L_HUGE ? (uint64_t)m : 0u;
    uint64_t x = own;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)x, o);
        const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(x >> 32), o);
        if (lane >= o) x += ((uint64_t)hi << 32) | lo;
    }
    if (lane == 63) whuge[w] = x;
    __syncthreads();
    if (threadIdx.x < 4) {
        unsigned tot = 0;
        for (int v = 0; v < W; ++v) tot += wcnt[threadIdx.x][v];
        bbase[threadIdx.x] = tot ? atomicAdd(&counts[threadIdx.x], tot) : 0u;
    } else if (threadIdx.x == 4) {
        unsigned long long tot = 0;
        for (int v = 0; v < W; ++v) tot += whuge[v];
        hbase = tot ? atomicAdd(huge_total, tot) : 0ull;
    }
    __syncthreads();
    if (cls >= 0) {
        unsigned before = 0;
        unsigned long long hbefore = 0;
        for (int v = 0; v < w; ++v) {
            before += wcnt[cls][v];
            hbefore += whuge[v];
        }
        const unsigned slot = bbase[cls] + before + (unsigned)__popcll(mk[cls] & lanemask_lt());
        lists[(size_t)cls * nq + slot] = g;
        if (cls == SEG_L_HUGE) huge_tmp[slot] = (int64_t)(hbase + hbefore + x - own);
    }
}

// segments of 2..64 hits, one wavefront each (a persistent grid over the
// list): lane l reads slot l, a wave scan of the slots' element counts gives
// each slot's first element, and lane e finds the slot of element e by a
// binary search over the lanes' offsets
__global__ __launch_bounds__(256) void segsort_small_kernel(const int64_t* __restrict__ off,
                                                            const int* __restrict__ list,
                                                            const unsigned* __restrict__ count,
                                                            const uint32_t* __restrict__ used,
                                                            const float* __restrict__ qxyz,
                                                            const float* __restrict__ radii, float r_all,
                                                            const float4* __restrict__ pts, uint32_t npts,
                                                            uint32_t* ids, int64_t capacity) {
    const int lane = threadIdx.x & 63;
    const unsigned n = *count, nw = (gridDim.x * blockDim.x) >> 6;
    for (unsigned li = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; li < n; li += nw) {
        const int seg = list[li];
        int64_t b, e;
        seg_bounds(off, seg, capacity, b, e);
        const int m = (int)(e - b);
        const int u = (int)min(used[seg], (uint32_t)m);
        const SegQ q = seg_query(qxyz, radii, r_all, seg);
        const uint32_t v = lane < u ? ids[b + lane] : RUN_LEN;
        const uint32_t nx = __shfl_down(v, 1);
        const uint32_t c = lane < u ? slot_elems(v, lane + 1 < u ? nx : 0u) : 0u;
        uint32_t x = c;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        const uint32_t ex = x - c;
        int lo = 0;
        for (int step = 32; step > 0; step >>= 1)
            if (__shfl(ex, lo + step) <= (uint32_t)lane) lo += step;
        const uint32_t sv = __shfl(v, lo), so = __shfl(ex, lo);
        const uint32_t p0 = (sv & RUN_START) ? (sv & SLOT_MASK) + ((uint32_t)lane - so) : (sv & SLOT_MASK);
        double d = INFINITY;
        uint32_t id = 0xFFFFFFFFu;
        if (lane < m) {
            const float4 p = pts[min(p0, npts - 1)];
            d = d2d(q.x, q.y, q.z, p.x, p.y, p.z);
            id = (uint32_t)__float_as_int(p.w);
        }
        for (int size = 2; size <= 64; size <<= 1)
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                const double od = __shfl_xor(d, stride);
                const uint32_t oi = __shfl_xor(id, stride);
                const bool up = (lane & size) == 0;
                const bool lower = (lane & stride) == 0;
                // keep the smaller element on the lower lane of an ascending pair
                const bool take_other = lower == up ? pair_gt(d, id, od, oi) : pair_gt(od, oi, d, id);
                if (take_other) { d = od; id = oi; }
            }
        if (lane < m) ids[b + lane] = id;
    }
}

// exclusive prefix sum of a[0..n) in place by the whole block of T lanes
template <int T>
__device__ __forceinline__ void block_exscan(uint32_t* a, int n, uint32_t* wsum) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int per = (n + T - 1) / T;
    const int b = min(t * per, n), e = min(b + per, n);
    uint32_t s = 0;
    for (int j = b; j < e; ++j) s += a[j];
    uint32_t x = s;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t before = 0;
    for (int v = 0; v < w; ++v) before += wsum[v];
    uint32_t run = before + x - s;
    for (int j = b; j < e; ++j) {
        const uint32_t v = a[j];
        a[j] = run;
        run += v;
    }
}

__device__ __forceinline__ int seg_bucket(double d, double scale, int nb) {
    const double f = d * scale;
    return f < (double)(nb - 1) ? (int)f : nb - 1;
}

// Expands a range's slots[0..u) into element positions in walk order, calling
// emit(e, position) for each element e < limit; returns the number of elements
// emitted.  T slots at a time: a block scan of their element counts, then
// element e of the batch lies in the last slot whose first element is <= e (a
// binary search over the batch's offsets in LDS).  offs / vals: LDS scratch of
// T words each; wsum: T / 64 words.  Called by the whole block; ends
// synchronised.
template <int T, class Emit>
__device__ int expand_slots(const uint32_t* __restrict__ slots, int u, int limit, uint32_t* offs, uint32_t* vals,
                            uint32_t* wsum, Emit emit) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int base = 0;
    for (int s0 = 0; s0 < u; s0 += T) {
        const int j = s0 + t;
        const uint32_t v = j < u ? slots[j] : RUN_LEN;
        const uint32_t c = j < u ? slot_elems(v, j + 1 < u ? slots[j + 1] : 0u) : 0u;
        uint32_t x = c;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int k = 0; k < T / 64; ++k) {
            const uint32_t ws = wsum[k];
            before += k < w ? ws : 0u;
            total += ws;
        }
        offs[t] = before + x - c;
        vals[t] = v;
        __syncthreads();
        const int nb = (int)min(total, (uint32_t)(limit - base));
        for (int e = t; e < nb; e += T) {
            int lo = 0;
            for (int step = T / 2; step > 0; step >>= 1)
                if (offs[lo + step] <= (uint32_t)e) lo += step;
            const uint32_t sv = vals[lo];
            emit(base + e, (sv & RUN_START) ? (sv & SLOT_MASK) + ((uint32_t)e - offs[lo]) : (sv & SLOT_MASK));
        }
        base += nb;
        __syncthreads();
    }
    return base;
}

template <int CAP, int T>
struct SegRankLds {
    double d[CAP];     // the expanded positions (as uint32), then the keys in bucket order
    uint32_t id[CAP];
    uint32_t cnt[CAP / 2 + 1];  // bucket counts, then bucket ends; first the expansion scratch
    uint32_t wsum[T / 64];
    int flag;
};

// Sorts the m <= CAP keys (dv, iv)[i] of lane t (element t + i*T) by (d64, id)
// and calls out(j, d64, id) once for the element of sorted position j.  All
// d64 <= r2.  Bucket counts, an exclusive scan, the keys scattered into bucket
// order in LDS, then each element's rank among the smaller keys of its bucket;
// a bucket of more than SEG_RUN elements switches to a bitonic network over the
// bucket-ordered keys.  Called by the whole block, which must be synchronised
// on entry (S.cnt and S.d / S.id free); ends synchronised.
template <int CAP, int T, class Out>
__device__ void rank_sort(SegRankLds<CAP, T>& S, int m, double r2, const double* dv, const uint32_t* iv, Out out) {
    constexpr int PER = CAP / T;
    const int t = threadIdx.x;
    const int nb = (m + 1) / 2;
    const double scale = r2 > 0.0 ? (double)nb / r2 : 0.0;
    for (int j = t; j <= nb; j += T) S.cnt[j] = 0;
    if (t == 0) S.flag = 0;
    int bk[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) bk[i] = t + i * T < m ? seg_bucket(dv[i], scale, nb) : 0;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i)
        if (t + i * T < m) atomicAdd(&S.cnt[bk[i]], 1u);
    __syncthreads();
    block_exscan<T>(S.cnt, nb, S.wsum);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i)
        if (t + i * T < m) {
            const uint32_t slot = atomicAdd(&S.cnt[bk[i]], 1u);
            S.d[slot] = dv[i];
            S.id[slot] = iv[i];
        }
    __syncthreads();
    // S.cnt[k] is now the end of bucket k (the start of bucket k + 1)
    bool big = false;
#pragma unroll
    for (int i = 0; i < PER; ++i)
        if (t + i * T < m) {
            const int st = bk[i] ? (int)S.cnt[bk[i] - 1] : 0, en = (int)S.cnt[bk[i]];
            big |= en - st > SEG_RUN;
        }
    if (big) S.flag = 1;
    __syncthreads();
    if (!S.flag) {
        // keys are distinct (ids differ), so the ranks place every element once
#pragma unroll
        for (int i = 0; i < PER; ++i)
            if (t + i * T < m) {
                const int st = bk[i] ? (int)S.cnt[bk[i] - 1] : 0, en = (int)S.cnt[bk[i]];
                int r = 0;
                for (int y = st; y < en; ++y) r += pair_gt(dv[i], iv[i], S.d[y], S.id[y]);
                out(st + r, dv[i], iv[i]);
            }
    } else {
        // many equal distances: a bitonic network over all the keys
        int Pn = 1;
        while (Pn < m) Pn <<= 1;
        for (int j = m + t; j < Pn; j += T) {
            S.d[j] = INFINITY;
            S.id[j] = 0xFFFFFFFFu;
        }
        __syncthreads();
        for (int size = 2; size <= Pn; size <<= 1)
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                for (int u = t; u < Pn / 2; u += T) {
                    const int j = 2 * u - (u & (stride - 1)), pj = j + stride;
                    const bool up = (j & size) == 0;
                    const double da = S.d[j], dc = S.d[pj];
                    const uint32_t ia = S.id[j], ic = S.id[pj];
                    if (pair_gt(da, ia, dc, ic) == up) {
                        S.d[j] = dc;
                        S.id[j] = ic;
                        S.d[pj] = da;
                        S.id[pj] = ia;
                    }
                }
                __syncthreads();
            }
        for (int j = t; j < m; j += T) out(j, S.d[j], S.id[j]);
    }
    __syncthreads();
}

// one workgroup per listed segment of up to CAP hits (a persistent grid
// striding over the list): positions expanded, points gathered into
// registers, bucket counts, scatter of the keys into bucket order, then each
// element's rank inside its bucket gives its output slot
template <int CAP, int T>
__global__ __launch_bounds__(T) void segsort_rank_kernel(const int64_t* __restrict__ off,
                                                         const int* __restrict__ list,
                                                         const unsigned* __restrict__ count,
                                                         const uint32_t* __restrict__ used,
                                                         const float* __restrict__ qxyz,
                                                         const float* __restrict__ radii, float r_all,
                                                         const float4* __restrict__ pts, uint32_t npts,
                                                         uint32_t* ids, int64_t capacity) {
    static_assert(CAP / 2 + 1 >= 2 * T, "expansion scratch lives in S.cnt");
    static_assert(CAP % T == 0, "elements per lane");
    constexpr int PER = CAP / T;
    __shared__ SegRankLds<CAP, T> S;
    const int t = threadIdx.x;
    const unsigned n = *count;
    uint32_t* P = reinterpret_cast<uint32_t*>(S.d);
    for (unsigned li = blockIdx.x; li < n; li += gridDim.x) {
        const int g = list[li];
        int64_t b, e;
        seg_bounds(off, g, capacity, b, e);
        const int m = (int)(e - b);
        const SegQ q = seg_query(qxyz, radii, r_all, g);
        // positions of the segment's hits (ends synchronised)
        const int ne = expand_slots<T>(ids + b, (int)min(used[g], (uint32_t)m), m, S.cnt, S.cnt + T, S.wsum,
                                       [&](int x, uint32_t p) { P[x] = p; });
        uint32_t pos[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int x = t + i * T;
            pos[i] = x < ne ? min(P[x], npts - 1) : 0u;
        }
        __syncthreads();  // P (over S.d) read
        double dv[PER];
        uint32_t iv[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            dv[i] = INFINITY;
            iv[i] = 0xFFFFFFFFu;
            if (t + i * T < m) {
                const float4 p = pts[pos[i]];
                dv[i] = d2d(q.x, q.y, q.z, p.x, p.y, p.z);
                iv[i] = (uint32_t)__float_as_int(p.w);
            }
        }
        rank_sort<CAP, T>(S, m, q.r2, dv, iv, [&](int j, double, uint32_t id) { ids[b + j] = id; });
    }
}

// Segments of CAP + 1 .. LIM hits (LIM <= 2 CAP), one workgroup each (a
// persistent grid over the list of longer segments): positions expanded into
// LDS (2 CAP of them fit over S.d), keys gathered into registers (LIM / T per
// lane), one
// bucket count over CAP / 2 buckets by distance and its scan, then the buckets
// in consecutive parts of at most CAP elements: each part's keys scattered into
// LDS in bucket order, ranked inside their buckets and written (a bitonic
// network over the part when one of its buckets holds more than SEG_RUN).  The
// parts cover increasing distance ranges, so sorting each part sorts the
// segment.  A longer segment, or one with a single bucket of more than CAP
// elements, is appended untouched to the retry list (with its merge-scratch
// offset) for segsort_huge_kernel.
template <int CAP, int T, int LIM>
__global__ __launch_bounds__(T) void segsort_parts_kernel(const int64_t* __restrict__ off,
                                                          const int* __restrict__ list,
                                                          const unsigned* __restrict__ count,
                                                          const uint32_t* __restrict__ used,
                                                          const float* __restrict__ qxyz,
                                                          const float* __restrict__ radii, float r_all,
                                                          const float4* __restrict__ pts, uint32_t npts,
                                                          uint32_t* ids, int64_t capacity,
                                                          const int64_t* __restrict__ huge_tmp, int* retry_list,
                                                          int64_t* retry_tmp, unsigned* retry_count) {
    static_assert(CAP / 2 + 1 >= 2 * T, "expansion scratch lives in S.cnt");
    static_assert(LIM % T == 0 && LIM <= 2 * CAP, "elements per lane");
    constexpr int PER = LIM / T;
    constexpr int NB = CAP / 2;  // buckets
    __shared__ SegRankLds<CAP, T> S;
    __shared__ int part_end[NB + 2];  // end buckets of the parts; [NB + 1]: their number
    const int t = threadIdx.x;
    const unsigned n = *count;
    uint32_t* P = reinterpret_cast<uint32_t*>(S.d);
    for (unsigned li = blockIdx.x; li < n; li += gridDim.x) {
        const int g = list[li];
        int64_t b, e;
        seg_bounds(off, g, capacity, b, e);
        const int m = (int)(e - b);
        if (m > LIM) {
            if (t == 0) {
                const unsigned j = atomicAdd(retry_count, 1u);
                retry_list[j] = g;
                retry_tmp[j] = huge_tmp[li];
            }
            continue;
        }
        const SegQ q = seg_query(qxyz, radii, r_all, g);
        const int ne = expand_slots<T>(ids + b, (int)min(used[g], (uint32_t)m), m, S.cnt, S.cnt + T, S.wsum,
                                       [&](int x, uint32_t p) { P[x] = p; });
        uint32_t pos[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int x = t + i * T;
            pos[i] = x < ne ? min(P[x], npts - 1) : 0u;
        }
        __syncthreads();  // P (over S.d) read
        double dv[PER];
        uint32_t iv[PER];
        const double scale = q.r2 > 0.0 ? (double)NB / q.r2 : 0.0;
        // an element's bucket is recomputed where needed (registers are short
        // with 2 CAP / T keys per lane)
        auto bkt = [&](int i) { return seg_bucket(dv[i], scale, NB); };
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            dv[i] = INFINITY;
            iv[i] = 0xFFFFFFFFu;
            if (t + i * T < m) {
                const float4 p = pts[pos[i]];
                dv[i] = d2d(q.x, q.y, q.z, p.x, p.y, p.z);
                iv[i] = (uint32_t)__float_as_int(p.w);
            }
        }
        for (int j = t; j <= NB; j += T) S.cnt[j] = 0;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PER; ++i)
            if (t + i * T < m) atomicAdd(&S.cnt[bkt(i)], 1u);
        __syncthreads();
        block_exscan<T>(S.cnt, NB, S.wsum);
        __syncthreads();
        if (t == 0) {
            // S.cnt[k]: start of bucket k, S.cnt[NB] = m.  Parts: maximal runs
            // of buckets holding at most CAP elements; 0 parts if a bucket alone
            // holds more
            S.cnt[NB] = (uint32_t)m;
            int np = 0, b1 = 0;
            while (b1 < NB) {
                int lo = b1, hi = NB;  // largest b2 in [b1, NB] with start(b2) - start(b1) <= CAP
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (S.cnt[mid] - S.cnt[b1] <= (uint32_t)CAP) lo = mid;
                    else hi = mid - 1;
                }
                if (lo == b1) {
                    np = 0;
                    break;
                }
                part_end[np++] = lo;
                b1 = lo;
            }
            part_end[NB + 1] = np;
        }
        __syncthreads();
        const int np = part_end[NB + 1];
        if (np == 0) {
            if (t == 0) {
                const unsigned j = atomicAdd(retry_count, 1u);
                retry_list[j] = g;
                retry_tmp[j] = huge_tmp[li];
            }
            __syncthreads();
            continue;
        }
        int b1 = 0;
        for (int pi = 0; pi < np; ++pi) {
            const int b2 = part_end[pi];
            const uint32_t base = S.cnt[b1];  // still the start of bucket b1
            const int psize = (int)(S.cnt[b2] - base);
            if (t == 0) S.flag = 0;
            __syncthreads();  // every lane read base and psize before the scatter moves the counts
#pragma unroll
            for (int i = 0; i < PER; ++i)
                if (t + i * T < m && bkt(i) >= b1 && bkt(i) < b2) {
                    const uint32_t slot = atomicAdd(&S.cnt[bkt(i)], 1u) - base;
                    S.d[slot] = dv[i];
                    S.id[slot] = iv[i];
                }
            __syncthreads();
            // S.cnt[k] is now the end of bucket k for the part's buckets, and
            // S.cnt[b1 - 1] the end of the previous part's last bucket
            bool big = false;
#pragma unroll
            for (int i = 0; i < PER; ++i)
                if (t + i * T < m && bkt(i) >= b1 && bkt(i) < b2) {
                    const int bi = bkt(i);
                    const int st = bi ? (int)S.cnt[bi - 1] : 0, en = (int)S.cnt[bi];
                    big |= en - st > SEG_RUN;
                }
            if (big) S.flag = 1;
            __syncthreads();
            if (!S.flag) {
#pragma unroll
                for (int i = 0; i < PER; ++i)
                    if (t + i * T < m && bkt(i) >= b1 && bkt(i) < b2) {
                        const int bi = bkt(i);
                        const int st = bi ? (int)S.cnt[bi - 1] : 0, en = (int)S.cnt[bi];
                        int r = 0;
                        for (int y = st; y < en; ++y)
                            r += pair_gt(dv[i], iv[i], S.d[y - (int)base], S.id[y - (int)base]);
                        ids[b + st + r] = iv[i];
                    }
            } else {
                int Pn = 1;
                while (Pn < psize) Pn <<= 1;
                for (int j = psize + t; j < Pn; j += T) {
                    S.d[j] = INFINITY;
                    S.id[j] = 0xFFFFFFFFu;
                }
                __syncthreads();
                for (int size = 2; size <= Pn; size <<= 1)
                    for (int stride = size >> 1; stride > 0; stride >>= 1) {
                        for (int u = t; u < Pn / 2; u += T) {
                            const int j = 2 * u - (u & (stride - 1)), pj = j + stride;
                            const bool up = (j & size) == 0;
                            const double da = S.d[j], dc = S.d[pj];
                            const uint32_t ia = S.id[j], ic = S.id[pj];
                            if (pair_gt(da, ia, dc, ic) == up) {
                                S.d[j] = dc;
                                S.id[j] = ic;
                                S.d[pj] = da;
                                S.id[pj] = ia;
                            }
                        }
                        __syncthreads();
                    }
                for (int j = t; j < psize; j += T) ids[b + base + j] = S.id[j];
            }
            __syncthreads();
            b1 = b2;
        }
    }
}

// number of elements of the sorted run (rd, ri)[0..n) that order before (d, i)
__device__ __forceinline__ int rank_in(const double* __restrict__ rd, const uint32_t* __restrict__ ri, int n,
                                       double d, uint32_t i) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (pair_gt(d, i, rd[mid], ri[mid])) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One workgroup per segment of more than SEG_LARGE hits: the slots expanded
// into positions in the merge scratch A, SEG_LARGE-element chunks of them
// bucket-sorted in LDS back into A, then runs merged
// pairwise (A -> B -> A ...); each element finds its output slot as its index
// in its own run plus its rank in the partner run ((d64, id) keys are distinct
// within a segment, so ranks never collide).  The last pass writes the ids.
constexpr int SEG_HUGE_T = 1024;
__global__ __launch_bounds__(SEG_HUGE_T) void segsort_huge_kernel(const int64_t* __restrict__ off,
                                                                  const int* __restrict__ list,
                                                                  const int64_t* __restrict__ huge_tmp,
                                                                  const unsigned* count,
                                                                  const uint32_t* __restrict__ used,
                                                                  const float* __restrict__ qxyz,
                                                                  const float* __restrict__ radii, float r_all,
                                                                  const float4* __restrict__ pts, uint32_t npts,
                                                                  uint32_t* ids, int64_t capacity, double* tmp_d,
                                                                  uint32_t* tmp_i, int64_t tmp_half) {
    __shared__ SegRankLds<SEG_LARGE, SEG_HUGE_T> S;
    if (blockIdx.x >= *count) return;
    const int g = list[blockIdx.x];
    int64_t b, e;
    seg_bounds(off, g, capacity, b, e);
    const int m = (int)(e - b);
    const SegQ q = seg_query(qxyz, radii, r_all, g);
    double* ad = tmp_d + huge_tmp[blockIdx.x];
    uint32_t* ai = tmp_i + huge_tmp[blockIdx.x];
    double* bd = ad + tmp_half;
    uint32_t* bi = ai + tmp_half;
    const int nexp = expand_slots<SEG_HUGE_T>(ids + b, (int)min(used[g], (uint32_t)m), m, S.cnt,
                                              S.cnt + SEG_HUGE_T, S.wsum,
                                              [&](int e, uint32_t p) { ai[e] = min(p, npts - 1); });
    for (int e = nexp + threadIdx.x; e < m; e += SEG_HUGE_T) ai[e] = 0u;
    __syncthreads();
    constexpr int PER = SEG_LARGE / SEG_HUGE_T;
    for (int c0 = 0; c0 < m; c0 += SEG_LARGE) {
        const int cm = min(SEG_LARGE, m - c0);
        double dv[PER];
        uint32_t iv[PER];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int x = (int)threadIdx.x + i * SEG_HUGE_T;
            dv[i] = INFINITY;
            iv[i] = 0xFFFFFFFFu;
            if (x < cm) {
                const float4 p = pts[ai[c0 + x]];
                dv[i] = d2d(q.x, q.y, q.z, p.x, p.y, p.z);
                iv[i] = (uint32_t)__float_as_int(p.w);
            }
        }
        __syncthreads();  // the chunk's positions are read before its sorted keys replace them
        rank_sort<SEG_LARGE, SEG_HUGE_T>(S, cm, q.r2, dv, iv, [&](int j, double d, uint32_t id) {
            ad[c0 + j] = d;
            ai[c0 + j] = id;
        });
    }
    for (int w = SEG_LARGE; w < m; w <<= 1) {
        for (int j = threadIdx.x; j < m; j += SEG_HUGE_T) {
            const int run = j / w, pos = j - run * w;
            const int pbase = (run ^ 1) * w;
            const int base = min(run, run ^ 1) * w;
            const int pn = pbase < m ? min(w, m - pbase) : 0;
            const int r = pn ? rank_in(ad + pbase, ai + pbase, pn, ad[j], ai[j]) : 0;
            bd[base + pos + r] = ad[j];
            bi[base + pos + r] = ai[j];
        }
        __syncthreads();
        double* td = ad; ad = bd; bd = td;
        uint32_t* ti = ai; ai = bi; bi = ti;
    }
    for (int j = threadIdx.x; j < m; j += SEG_HUGE_T) ids[b + j] = ai[j];
}

template <class K>
static int persistent_blocks(K kernel, int threads, int num_cus, std::atomic<int>* cache) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    int g = cache[dev & 63].load(std::memory_order_relaxed);
    if (g > 0) return g;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    g = per_cu * num_cus;
    cache[dev & 63].store(g, std::memory_order_relaxed);
    return g;
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
int range_run(Ctx* ctx, const Tree& T, const float* d_q, int nq, float r, int64_t* d_off, uint32_t* d_ids,
              int64_t capacity, DevBuf& counts_buf, DevBuf& d2_buf, DevBuf& big_buf, DevBuf& keys_buf,
              DevBuf& perm_buf, DevBuf& qs4_buf, DevBuf& sums_buf, int64_t* h_total, const float* d_radii,
              bool count, bool time_walk) {
    (void)d2_buf;
    ctx->kernel_timed = false;
    hipStream_t s = ctx->stream;
    unsigned* counters = ctx->counters.as<unsigned>();
    unsigned long long* stats64 = (unsigned long long*)(counters + CNT_TOTAL);
    GPRT_TRY(keys_buf.reserve(8ull * nq));
    GPRT_TRY(perm_buf.reserve(std::max<size_t>(4ull * nq, query_prep_scratch_bytes(1, 1))));
    GPRT_TRY(qs4_buf.reserve(16ull * nq));
    GPRT_TRY(counts_buf.reserve(8ull * nq + 8));
    const int tiles = (nq + SCAN_TILE - 1) / SCAN_TILE;
    GPRT_TRY(sums_buf.reserve(8ull * tiles + 64));
    GPRT_TRY(ctx->ws[7].reserve(64));
    int64_t* qoff = (int64_t*)(ctx->ws[7].as<char>() + 32);
    int64_t h_qoff[2] = {0, nq};
    GPRT_HIP_TRY(hipMemcpyAsync(qoff, h_qoff, 16, hipMemcpyHostToDevice, s));
    GPRT_TRY(launch_fill_u32(s, counters, CNT_TOTAL + 2 * CNT_STATS64, 0u));
    GPRT_TRY(launch_query_prep(ctx, d_q, qoff, 1, T.key_bbox.as<float>(), 1, nq, keys_buf.as<uint64_t>(),
                               perm_buf.as<uint32_t>(), qs4_buf.as<float4>(), counters, nq, nullptr,
                               perm_buf.as<unsigned>(), counters + CNT_TOTAL - 1));
    int2 root;
    GPRT_HIP_TRY(hipMemcpyAsync(&root, T.roots.ptr, sizeof(int2), hipMemcpyDeviceToHost, s));
    GPRT_HIP_TRY(hipStreamSynchronize(s));
    const int grid = (nq + 255) / 256;
    const int max_steps = T.n + 64;
    const Node* nodes = T.nodes.as<Node>();
    const float4* pts = T.pts4.as<float4>();
    const float4* qs4 = qs4_buf.as<float4>();
    if (time_walk) GPRT_HIP_TRY(hipEventRecord(ctx->ev2, s));
    if (!d_ids) {
        if (count)
            hipLaunchKernelGGL((range_kernel<false, true>), dim3(grid), dim3(256), 0, s, nodes, pts, root, qs4, nq, r,
                               d_radii, counts_buf.as<int64_t>(), nullptr, nullptr, (int64_t)0, nullptr, max_steps,
                               counters + CNT_ERROR, counters + CNT_OFFSETS, stats64);
        else
            hipLaunchKernelGGL((range_kernel<false, false>), dim3(grid), dim3(256), 0, s, nodes, pts, root, qs4, nq,
                               r, d_radii, counts_buf.as<int64_t>(), nullptr, nullptr, (int64_t)0, nullptr, max_steps,
                               counters + CNT_ERROR, counters + CNT_OFFSETS, stats64);
        if (time_walk) {
            GPRT_HIP_TRY(hipEventRecord(ctx->ev3, s));
            ctx->kernel_timed = true;
        }
        int64_t* sums = sums_buf.as<int64_t>();
        int64_t* total = sums + tiles;
        hipLaunchKernelGGL(scan_tile_sums, dim3(tiles), dim3(256), 0, s, counts_buf.as<int64_t>(), nq, sums);
        hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(1024), 0, s, sums, tiles, total);
        hipLaunchKernelGGL(scan_tile_apply, dim3(tiles), dim3(256), 0, s, counts_buf.as<int64_t>(), nq, sums, d_off);
        hipLaunchKernelGGL(set_last_kernel, dim3(1), dim3(1), 0, s, d_off, nq, total);
        GPRT_HIP_TRY(hipGetLastError());
        GPRT_HIP_TRY(hipMemcpyAsync(h_total, total, 8, hipMemcpyDeviceToHost, s));
        GPRT_HIP_TRY(hipStreamSynchronize(s));
        return GPRT_OK;
    }
    // segment lists (4 x nq ints), the long segments' merge-scratch offsets
    // (nq int64), the retry list (nq ints) and its offsets (nq int64)
    GPRT_TRY(big_buf.reserve(16ull * nq + 8ull * nq + 4ull * (nq + 1) + 8ull * nq + 16));
    int* lists = big_buf.as<int>();
    int64_t* huge_tmp = (int64_t*)(lists + 4 * (size_t)nq);
    // slots written per query, in the count buffer (unused by the fill call)
    uint32_t* used = counts_buf.as<uint32_t>();
    if (count)
        hipLaunchKernelGGL((range_kernel<true, true>), dim3(grid), dim3(256), 0, s, nodes, pts, root, qs4, nq, r,
                           d_radii, nullptr, d_off, d_ids, capacity, used, max_steps, counters + CNT_ERROR,
                           counters + CNT_OFFSETS, stats64);
    else
        hipLaunchKernelGGL((range_kernel<true, false>), dim3(grid), dim3(256), 0, s, nodes, pts, root, qs4, nq, r,
                           d_radii, nullptr, d_off, d_ids, capacity, used, max_steps, counters + CNT_ERROR,
                           counters + CNT_OFFSETS, stats64);
    if (time_walk) {
        GPRT_HIP_TRY(hipEventRecord(ctx->ev3, s));
        ctx->kernel_timed = true;
    }
    // list lengths in counters the query sort does not touch (cleared above);
    // stats64[3]: the huge segments' total
    unsigned* cnt = counters + CNT_AUX;
    unsigned long long* huge_total = stats64 + 3;
    const uint32_t npts = (uint32_t)T.n;
    hipLaunchKernelGGL(seg_classify_kernel, dim3((nq + CLS_T - 1) / CLS_T), dim3(CLS_T), 0, s, d_off, nq, capacity, pts,
                       npts, d_ids, lists,
                       cnt, huge_tmp, huge_total);
    static std::atomic<int> g_small[64], g_mid[64], g_large[64];
    hipLaunchKernelGGL(segsort_small_kernel, dim3(persistent_blocks(segsort_small_kernel, 256, ctx->num_cus, g_small)),
                       dim3(256), 0, s, d_off, lists + (size_t)SEG_L_SMALL * nq, cnt + SEG_L_SMALL, used, d_q,
                       d_radii, r, pts, npts, d_ids, capacity);
    hipLaunchKernelGGL((segsort_rank_kernel<SEG_MID, 256>),
                       dim3(persistent_blocks(segsort_rank_kernel<SEG_MID, 256>, 256, ctx->num_cus, g_mid)),
                       dim3(256), 0, s, d_off, lists + (size_t)SEG_L_MID * nq, cnt + SEG_L_MID, used, d_q, d_radii, r,
                       pts, npts, d_ids, capacity);
    hipLaunchKernelGGL((segsort_rank_kernel<SEG_LARGE, 1024>),
                       dim3(persistent_blocks(segsort_rank_kernel<SEG_LARGE, 1024>, 1024, ctx->num_cus, g_large)),
                       dim3(1024), 0, s, d_off, lists + (size_t)SEG_L_LARGE * nq, cnt + SEG_L_LARGE, used, d_q,
                       d_radii, r, pts, npts, d_ids, capacity);
    static std::atomic<int> g_parts[64];
    int* retry_list = reinterpret_cast<int*>(huge_tmp + nq);
    int64_t* retry_tmp = huge_tmp + nq + (nq + 1) / 2;
    unsigned* retry_count = counters + CNT_RETRY;
    hipLaunchKernelGGL((segsort_parts_kernel<SEG_LARGE, 1024, SEG_PARTS>),
                       dim3(persistent_blocks(segsort_parts_kernel<SEG_LARGE, 1024, SEG_PARTS>, 1024, ctx->num_cus,
                                              g_parts)),
                       dim3(1024), 0, s, d_off, lists + (size_t)SEG_L_HUGE * nq, cnt + SEG_L_HUGE, used, d_q, d_radii,
                       r, pts, npts, d_ids, capacity, huge_tmp, retry_list, retry_tmp, retry_count);
    struct {
        unsigned nretry;
        unsigned long long total;
    } h{0, 0};
    GPRT_HIP_TRY(hipMemcpyAsync(&h.nretry, retry_count, 4, hipMemcpyDeviceToHost, s));
    GPRT_HIP_TRY(hipMemcpyAsync(&h.total, huge_total, 8, hipMemcpyDeviceToHost, s));
    GPRT_HIP_TRY(hipStreamSynchronize(s));
    if (h.nretry) {
        // the merge scratch spans every long segment's offset (from the classify pass)
        GPRT_TRY(ctx->sort_tmp[0].reserve(16ull * h.total));
        GPRT_TRY(ctx->sort_tmp[1].reserve(8ull * h.total));
        hipLaunchKernelGGL(segsort_huge_kernel, dim3(h.nretry), dim3(SEG_HUGE_T), 0, s, d_off, retry_list, retry_tmp,
                           retry_count, used, d_q, d_radii, r, pts, npts, d_ids, capacity,
                           ctx->sort_tmp[0].as<double>(), ctx->sort_tmp[1].as<uint32_t>(), (int64_t)h.total);
    }
    GPRT_HIP_TRY(hipGetLastError());
    return GPRT_OK;
}

}  // namespace gprt
