// sah.hip -- the SAH tree build (GPRT_BUILD_SAH): the points in the LBVH's
// Morton order are split top-down by binned SAH (the decisions of sah.hpp) with
// stable partitions, so every subtree stays a contiguous range of the sorted
// points, in the node layout of the LBVH (child boxes and references, leaf
// buckets of at most leaf_size points, parent links and ranges for the
// bottom-up refit that then fills the boxes).
//
// Node ids: the Karras numbering (sah_seg_id: a cloud's root is its first
// point, an internal child A its last point, an internal child B its first
// point), so the build needs no allocator, the tree depends only on the points,
// and every walk finds a child's points as it does in the LBVH.
// Subtrees of at most leaf_size points are still split (at their index median)
// down to single points: the walks scan them as buckets, and the refit, like
// the Karras refit, aggregates over the full binary tree.
//
// Three phases:
//  * segments of more than SAH_SMALL points, one level at a time over the whole
//    GPU: chunk list, binning of each chunk in LDS merged by global atomics,
//    the split chosen per segment, stable scatter into the other buffers, copy
//    back (one host read per level: the segment and chunk counts);
//  * segments of leaf_size < m <= SAH_SMALL points: one wavefront each, with a
//    segment stack in LDS (the larger half pushed first, so at most log2(m)
//    entries), binning with LDS atomics and ballot partitions;
//  * subtrees of 2..leaf_size points: one thread each, index-median splits.
#include <algorithm>

#include "knn_internal.hpp"
#include "sah.hpp"

namespace gprt {

#ifndef GPRT_SAH_SMALL
#define GPRT_SAH_SMALL 1024
#endif
constexpr int SAH_SMALL = GPRT_SAH_SMALL;  // largest segment the wavefront phase takes
constexpr int SAH_CH = 2048;     // points per chunk of the large-segment passes
constexpr int SAH_CT = 256;      // threads of a chunk workgroup (8 points each)
constexpr int SAH_STACK = 24;    // segment stack of a wavefront (log2(1024) + 1 used at most)

// counters
constexpr int SC_CUR = 0, SC_NEXT = 1, SC_SMALL = 2, SC_BUCKET = 3, SC_CHUNKS = 4, SC_ERR = 5, SC_WORDS = 8;
constexpr int SAH_BBUF = 64;  // buckets a wavefront of the small-segment kernel batches before appending them

struct SahSeg {
    int lo, hi;
    int slot;         // (parent node << 1 | side), or -(cloud + 2) for a cloud's root
    int first_chunk;  // set by the chunk-list kernel
    float blo[3], bhi[3];  // exact box of the segment's points
    int nchunks, depth;
};
struct SahChunk {
    int seg, start, end, pad;
};
struct SahCut {
    int axis, bin;  // axis < 0: index-median split, points stay
    uint32_t nleft;
    int pad;
};

struct SahArgs {
    float4* pts;
    uint32_t* perm;
    uint64_t* codes;
    float4* apts;
    uint32_t* aperm;
    uint64_t* acodes;
    Node* nodes;
    int* parent;
    int2* ranges;
    int2* roots;
    int n, leaf;
    SahSeg* cur;
    SahSeg* next;
    SahSeg* small;
    int4* buckets;  // (lo, hi, slot)
    SahChunk* chunks;
    uint32_t* chunk_cnt;   // [chunk][3 * SAH_BINS] bin counts of the chunk
    uint32_t* chunk_loff;  // [chunk] left points in the earlier chunks of its segment
    uint32_t* bins;        // [segment of the level][SAH_WORDS]
    SahCut* cuts;          // [segment of the level]
    unsigned* ctr;
    int cap_seg, cap_small, cap_bucket, cap_chunks;
};

__device__ __forceinline__ int parent_link(int slot) { return slot >= 0 ? slot : -1; }

// a new segment [lo, hi) under `slot` goes to the list its size asks for
__device__ void sah_emit(const SahArgs& a, int lo, int hi, int slot, int depth, const float* blo, const float* bhi) {
    const int m = hi - lo;
    if (m <= 0) return;
    if (m == 1) {
        a.parent[(a.n - 1) + lo] = parent_link(slot);
        return;
    }
    if (m <= a.leaf) {
        const unsigned j = atomicAdd(&a.ctr[SC_BUCKET], 1u);
        if (j < (unsigned)a.cap_bucket) a.buckets[j] = make_int4(lo, hi, slot, 0);
        else atomicAdd(&a.ctr[SC_ERR], 1u);
        return;
    }
    SahSeg g;
    g.lo = lo;
    g.hi = hi;
    g.slot = slot;
    g.first_chunk = 0;
    g.nchunks = 0;
    g.depth = depth;
    for (int k = 0; k < 3; ++k) {
        g.blo[k] = blo[k];
        g.bhi[k] = bhi[k];
    }
    if (m > SAH_SMALL) {
        const unsigned j = atomicAdd(&a.ctr[SC_NEXT], 1u);
        if (j < (unsigned)a.cap_seg) a.next[j] = g;
        else atomicAdd(&a.ctr[SC_ERR], 1u);
    } else {
        const unsigned j = atomicAdd(&a.ctr[SC_SMALL], 1u);
        if (j < (unsigned)a.cap_small) a.small[j] = g;
        else atomicAdd(&a.ctr[SC_ERR], 1u);
    }
}

// The node of the split of [lo, hi) at mid (one thread): range, parent link,
// child counts and references (a half of at most leaf_size points is a
// bucket, a larger one the internal node mid - 1 (A) or mid (B)), and the
// cloud's root when the segment is a whole cloud larger than a bucket.
// Returns the node's id.
__device__ int sah_node(const SahArgs& a, int lo, int mid, int hi, int slot) {
    const int id = sah_seg_id(lo, hi, slot);
    a.ranges[id] = make_int2(lo, hi - 1);
    a.parent[id] = parent_link(slot);
    Node* nd = a.nodes + id;
    nd->ahi.w = __int_as_float(mid - lo);
    nd->bhi.w = __int_as_float(hi - mid);
    nd->alo.w = __int_as_float(mid - lo <= a.leaf ? ~lo : mid - 1);
    nd->blo.w = __int_as_float(hi - mid <= a.leaf ? ~mid : mid);
    if (hi - lo > a.leaf && slot < 0) a.roots[-slot - 2] = make_int2(id, hi - lo);
    return id;
}

// ids that no cloud owns (the last id of each cloud's range) become empty
// nodes, so the 4-wide view built over every id reads nothing past them
__global__ void sah_unused_kernel(SahArgs a, const int64_t* __restrict__ off, int nclouds) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nclouds) return;
    const int64_t id = off[c + 1] - 1;
    if (id < 0 || id >= a.n - 1) return;
    Node e;
    e.alo = e.blo = make_float4(0.f, 0.f, 0.f, __int_as_float(~0));
    e.ahi = e.bhi = make_float4(0.f, 0.f, 0.f, __int_as_float(0));
    a.nodes[id] = e;
    a.ranges[id] = make_int2(0, -1);
    a.parent[id] = -1;
}

// every cloud's points as the first segment (its box: the cloud's exact bounds)
__global__ void sah_roots_kernel(SahArgs a, const int64_t* __restrict__ off, const float* __restrict__ bbox,
                                 int nclouds) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nclouds) return;
    const float blo[3] = {bbox[6 * c], bbox[6 * c + 1], bbox[6 * c + 2]};
    const float bhi[3] = {bbox[6 * c + 3], bbox[6 * c + 4], bbox[6 * c + 5]};
    sah_emit(a, (int)off[c], (int)off[c + 1], -(c + 2), 0, blo, bhi);
}

// The level's segments (last level's `next`, now `cur`): their chunks, listed
// in order, and their bin words reset.  One workgroup.
__global__ __launch_bounds__(1024) void sah_chunklist_kernel(SahArgs a) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const unsigned nnext = a.ctr[SC_NEXT];
    const unsigned ncur = min(nnext, (unsigned)a.cap_seg);
    __syncthreads();
    if (t == 0) {
        if (nnext > ncur) atomicAdd(&a.ctr[SC_ERR], 1u);
        a.ctr[SC_CUR] = ncur;
        a.ctr[SC_NEXT] = 0;
        carry = 0;
    }
    __syncthreads();
    for (unsigned base = 0; base < ncur; base += 1024) {
        const unsigned i = base + t;
        const uint32_t c = i < ncur ? (uint32_t)((a.cur[i].hi - a.cur[i].lo + SAH_CH - 1) / SAH_CH) : 0u;
        uint32_t x = c;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wv] = x;
        __syncthreads();
        uint32_t before = carry;
        for (int v = 0; v < wv; ++v) before += wsum[v];
        const uint32_t first = before + x - c;
        if (i < ncur) {
            SahSeg& g = a.cur[i];
            g.first_chunk = (int)first;
            g.nchunks = (int)c;
            for (uint32_t k = 0; k < c; ++k) {
                const uint32_t ci = first + k;
                if (ci >= (uint32_t)a.cap_chunks) {
                    atomicAdd(&a.ctr[SC_ERR], 1u);
                    break;
                }
                SahChunk ch;
                ch.seg = (int)i;
                ch.start = g.lo + (int)k * SAH_CH;
                ch.end = min(g.hi, ch.start + SAH_CH);
                ch.pad = 0;
                a.chunks[ci] = ch;
            }
        }
        __syncthreads();
        if (t == 1023) carry = before + x;
        __syncthreads();
    }
    if (t == 0) a.ctr[SC_CHUNKS] = carry;
    for (size_t j = t; j < (size_t)ncur * SAH_WORDS; j += 1024) a.bins[j] = sah_init_word((int)(j % SAH_WORDS));
}

// bins of one chunk in LDS, merged into its segment's words; the chunk's own
// bin counts kept for the partition offsets
__global__ __launch_bounds__(SAH_CT) void sah_bin_kernel(SahArgs a) {
    __shared__ uint32_t w[SAH_WORDS];
    const SahChunk ch = a.chunks[blockIdx.x];
    const SahSeg g = a.cur[ch.seg];
    for (int j = threadIdx.x; j < SAH_WORDS; j += SAH_CT) w[j] = sah_init_word(j);
    __syncthreads();
    float sc[3];
    for (int k = 0; k < 3; ++k) sc[k] = sah_scale(g.blo[k], g.bhi[k]);
    for (int i = ch.start + threadIdx.x; i < ch.end; i += SAH_CT) {
        const float4 p = a.pts[i];
        const float v[3] = {p.x, p.y, p.z};
        const uint32_t e[3] = {sah_enc(p.x), sah_enc(p.y), sah_enc(p.z)};
        for (int ax = 0; ax < 3; ++ax) {
            const int b = sah_bin(v[ax], g.blo[ax], sc[ax]);
            atomicAdd(&w[sah_cnt_word(ax, b)], 1u);
            for (int k = 0; k < 3; ++k) {
                atomicMin(&w[sah_lo_word(ax, b, k)], e[k]);
                atomicMax(&w[sah_hi_word(ax, b, k)], e[k]);
            }
        }
    }
    __syncthreads();
    uint32_t* gw = a.bins + (size_t)ch.seg * SAH_WORDS;
    for (int j = threadIdx.x; j < SAH_WORDS; j += SAH_CT) {
        const uint32_t v = w[j];
        if (j < 3 * SAH_BINS) {
            a.chunk_cnt[(size_t)blockIdx.x * (3 * SAH_BINS) + j] = v;
            if (v) atomicAdd(&gw[j], v);
        } else if (j < 12 * SAH_BINS) {
            if (v != 0xFFFFFFFFu) atomicMin(&gw[j], v);
        } else if (v) {
            atomicMax(&gw[j], v);
        }
    }
}

// one wavefront per segment of the level: the split, the left offsets of its
// chunks, its node and its two halves
__global__ __launch_bounds__(64) void sah_split_kernel(SahArgs a) {
    __shared__ uint32_t w[SAH_WORDS];
    __shared__ SahSplit sp;
    __shared__ int found;
    const int si = blockIdx.x, t = threadIdx.x;
    const SahSeg g = a.cur[si];
    for (int j = t; j < SAH_WORDS; j += 64) w[j] = a.bins[(size_t)si * SAH_WORDS + j];
    __syncthreads();
    if (t == 0) found = (g.depth < SAH_MAX_DEPTH && sah_choose(w, sp)) ? 1 : 0;
    __syncthreads();
    const int m = g.hi - g.lo;
    // left points of each chunk, exclusive prefix over the segment's chunks
    const int nch = g.nchunks, per = (nch + 63) / 64;
    const int c0 = g.first_chunk + min(t * per, nch), c1 = g.first_chunk + min((t + 1) * per, nch);
    uint32_t mine = 0;
    if (found)
        for (int c = c0; c < c1; ++c)
            for (int b = 0; b <= sp.bin; ++b) mine += a.chunk_cnt[(size_t)c * (3 * SAH_BINS) + sah_cnt_word(sp.axis, b)];
    uint32_t x = mine;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o);
        if (t >= o) x += y;
    }
    uint32_t run = x - mine;
    for (int c = c0; c < c1; ++c) {
        a.chunk_loff[c] = run;
        if (found)
            for (int b = 0; b <= sp.bin; ++b) run += a.chunk_cnt[(size_t)c * (3 * SAH_BINS) + sah_cnt_word(sp.axis, b)];
    }
    const uint32_t total = __shfl(x, 63);
    if (t != 0) return;
    SahCut cut;
    cut.pad = 0;
    int mid;
    float llo[3], lhi[3], rlo[3], rhi[3];
    if (found) {
        if (total != sp.nleft) atomicAdd(&a.ctr[SC_ERR], 1u);  // chunk counts disagree with the merged bins
        cut.axis = sp.axis;
        cut.bin = sp.bin;
        cut.nleft = sp.nleft;
        mid = g.lo + (int)sp.nleft;
        for (int k = 0; k < 3; ++k) {
            llo[k] = sp.llo[k]; lhi[k] = sp.lhi[k];
            rlo[k] = sp.rlo[k]; rhi[k] = sp.rhi[k];
        }
    } else {
        cut.axis = -1;
        cut.bin = 0;
        cut.nleft = (uint32_t)(m / 2);
        mid = g.lo + m / 2;
        for (int k = 0; k < 3; ++k) {
            llo[k] = rlo[k] = g.blo[k];
            lhi[k] = rhi[k] = g.bhi[k];
        }
    }
    a.cuts[si] = cut;
    const int id = sah_node(a, g.lo, mid, g.hi, g.slot);
    sah_emit(a, g.lo, mid, id << 1, g.depth + 1, llo, lhi);
    sah_emit(a, mid, g.hi, (id << 1) | 1, g.depth + 1, rlo, rhi);
}

// stable partition of one chunk into the other buffers: left points of the
// segment first (in order), then the right ones
__global__ __launch_bounds__(SAH_CT) void sah_scatter_kernel(SahArgs a) {
    __shared__ uint32_t wl[SAH_CT / 64], wr[SAH_CT / 64];
    const SahChunk ch = a.chunks[blockIdx.x];
    const SahCut cut = a.cuts[ch.seg];
    if (cut.axis < 0) return;
    const SahSeg g = a.cur[ch.seg];
    const float lo = g.blo[cut.axis], sc = sah_scale(g.blo[cut.axis], g.bhi[cut.axis]);
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    constexpr int PER = SAH_CH / SAH_CT;
    const int i0 = ch.start + t * PER;
    bool left[PER];
    uint32_t nl = 0, nr = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int i = i0 + j;
        left[j] = false;
        if (i < ch.end) {
            const float4 p = a.pts[i];
            const float c = cut.axis == 0 ? p.x : (cut.axis == 1 ? p.y : p.z);
            left[j] = sah_bin(c, lo, sc) <= cut.bin;
            nl += left[j];
            nr += !left[j];
        }
    }
    uint32_t xl = nl, xr = nr;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t yl = __shfl_up(xl, o), yr = __shfl_up(xr, o);
        if (lane >= o) {
            xl += yl;
            xr += yr;
        }
    }
    if (lane == 63) {
        wl[wv] = xl;
        wr[wv] = xr;
    }
    __syncthreads();
    uint32_t bl = xl - nl, br = xr - nr;
    for (int v = 0; v < wv; ++v) {
        bl += wl[v];
        br += wr[v];
    }
    const uint32_t loff = a.chunk_loff[blockIdx.x];
    const uint32_t roff = (uint32_t)(ch.start - g.lo) - loff;
    for (int j = 0; j < PER; ++j) {
        const int i = i0 + j;
        if (i >= ch.end) break;
        const int pos = left[j] ? g.lo + (int)(loff + bl++) : g.lo + (int)(cut.nleft + roff + br++);
        if (pos < g.lo || pos >= g.hi) {  // counts inconsistent: fail the build, stay in bounds
            atomicAdd(&a.ctr[SC_ERR], 1u);
            continue;
        }
        a.apts[pos] = a.pts[i];
        a.aperm[pos] = a.perm[i];
        a.acodes[pos] = a.codes[i];
    }
}

__global__ __launch_bounds__(SAH_CT) void sah_copyback_kernel(SahArgs a) {
    const SahChunk ch = a.chunks[blockIdx.x];
    if (a.cuts[ch.seg].axis < 0) return;
    for (int i = ch.start + threadIdx.x; i < ch.end; i += SAH_CT) {
        a.pts[i] = a.apts[i];
        a.perm[i] = a.aperm[i];
        a.codes[i] = a.acodes[i];
    }
}

// Segments of leaf_size < m <= SAH_SMALL points: one wavefront each (four per
// workgroup), the whole subtree in one pass of a segment stack.
__global__ __launch_bounds__(256) void sah_small_kernel(SahArgs a) {
    __shared__ uint32_t bins[4][SAH_WORDS];
    __shared__ SahSeg stack[4][SAH_STACK];
    __shared__ SahSplit spl[4];
    __shared__ int fnd[4];
    // buckets of the wavefront's subtrees, appended to the global list in
    // batches: one returning atomic per batch instead of one per bucket
    // (returning atomics on one counter serialise)
    __shared__ int4 bbuf[4][SAH_BBUF];
    __shared__ int bn[4];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint32_t* w = bins[wv];
    SahSeg* st = stack[wv];
    const unsigned nsmall = min(a.ctr[SC_SMALL], (unsigned)a.cap_small);
    const uint64_t lt = lanemask_lt();
    if (lane == 0) bn[wv] = 0;
    // a half of at most leaf_size points (lane 0): a single point's parent
    // link, or a bucket into the wavefront's batch
    auto emit_small = [&](int lo, int hi, int slot) {
        if (hi - lo <= 0) return;
        if (hi - lo == 1) {
            a.parent[(a.n - 1) + lo] = parent_link(slot);
            return;
        }
        bbuf[wv][bn[wv]++] = make_int4(lo, hi, slot, 0);
    };
    // the batch to the global list (the whole wavefront)
    auto flush = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int cnt = bn[wv];
        unsigned base = 0;
        if (lane == 0 && cnt > 0) base = atomicAdd(&a.ctr[SC_BUCKET], (unsigned)cnt);
        base = __shfl(base, 0);
        for (int j = lane; j < cnt; j += 64) {
            if (base + j < (unsigned)a.cap_bucket) a.buckets[base + j] = bbuf[wv][j];
            else atomicAdd(&a.ctr[SC_ERR], 1u);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (lane == 0) bn[wv] = 0;
    };
    for (unsigned si = blockIdx.x * 4 + wv; si < nsmall; si += gridDim.x * 4) {
        if (lane == 0) st[0] = a.small[si];
        int sp = 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        while (sp > 0) {
            --sp;
            const SahSeg g = st[sp];
            __builtin_amdgcn_wave_barrier();
            const int m = g.hi - g.lo;  // > leaf_size (smaller segments never enter the stack)
            const bool use_sah = g.depth < SAH_MAX_DEPTH;
            for (int j = lane; j < SAH_WORDS; j += 64) w[j] = sah_init_word(j);
            __builtin_amdgcn_wave_barrier();
            float sc[3];
            for (int k = 0; k < 3; ++k) sc[k] = sah_scale(g.blo[k], g.bhi[k]);
            for (int i = g.lo + lane; use_sah && i < g.hi; i += 64) {
                const float4 p = a.pts[i];
                const float v[3] = {p.x, p.y, p.z};
                const uint32_t e[3] = {sah_enc(p.x), sah_enc(p.y), sah_enc(p.z)};
                for (int ax = 0; ax < 3; ++ax) {
                    const int b = sah_bin(v[ax], g.blo[ax], sc[ax]);
                    atomicAdd(&w[sah_cnt_word(ax, b)], 1u);
                    for (int k = 0; k < 3; ++k) {
                        atomicMin(&w[sah_lo_word(ax, b, k)], e[k]);
                        atomicMax(&w[sah_hi_word(ax, b, k)], e[k]);
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane == 0) fnd[wv] = (use_sah && sah_choose(w, spl[wv])) ? 1 : 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const bool found = fnd[wv] != 0;
            const SahSplit s = spl[wv];
            int mid = g.lo + m / 2;
            float llo[3], lhi[3], rlo[3], rhi[3];
            for (int k = 0; k < 3; ++k) {
                llo[k] = rlo[k] = g.blo[k];
                lhi[k] = rhi[k] = g.bhi[k];
            }
            if (found) {
                mid = g.lo + (int)s.nleft;
                for (int k = 0; k < 3; ++k) {
                    llo[k] = s.llo[k]; lhi[k] = s.lhi[k];
                    rlo[k] = s.rlo[k]; rhi[k] = s.rhi[k];
                }
                const float blo = g.blo[s.axis], bsc = sc[s.axis];
                uint32_t bl = 0, br = 0;
                for (int i0 = g.lo; i0 < g.hi; i0 += 64) {
                    const int i = i0 + lane;
                    const bool valid = i < g.hi;
                    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
                    bool left = false;
                    if (valid) {
                        p = a.pts[i];
                        const float c = s.axis == 0 ? p.x : (s.axis == 1 ? p.y : p.z);
                        left = sah_bin(c, blo, bsc) <= s.bin;
                    }
                    const uint64_t L = __ballot(valid && left), R = __ballot(valid && !left);
                    if (valid) {
                        const int pos = left ? g.lo + (int)(bl + __popcll(L & lt))
                                             : g.lo + (int)(s.nleft + br + __popcll(R & lt));
                        if (pos >= g.lo && pos < g.hi) {
                            a.apts[pos] = p;
                            a.aperm[pos] = a.perm[i];
                            a.acodes[pos] = a.codes[i];
                        } else {
                            atomicAdd(&a.ctr[SC_ERR], 1u);
                        }
                    }
                    bl += (uint32_t)__popcll(L);
                    br += (uint32_t)__popcll(R);
                }
                __threadfence_block();
                __builtin_amdgcn_wave_barrier();
                for (int i = g.lo + lane; i < g.hi; i += 64) {
                    a.pts[i] = a.apts[i];
                    a.perm[i] = a.aperm[i];
                    a.codes[i] = a.acodes[i];
                }
                __threadfence_block();
                __builtin_amdgcn_wave_barrier();
            }
            const int id = sah_seg_id(g.lo, g.hi, g.slot);
            if (lane == 0) {
                sah_node(a, g.lo, mid, g.hi, g.slot);
                // halves of at most leaf_size points: single points and buckets
                if (mid - g.lo <= a.leaf) emit_small(g.lo, mid, id << 1);
                if (g.hi - mid <= a.leaf) emit_small(mid, g.hi, (id << 1) | 1);
                // larger halves onto the stack, the larger one first
                SahSeg h[2];
                int nh = 0;
                const int lo2[2] = {g.lo, mid}, hi2[2] = {mid, g.hi};
                for (int side = 0; side < 2; ++side) {
                    if (hi2[side] - lo2[side] <= a.leaf) continue;
                    SahSeg c;
                    c.lo = lo2[side];
                    c.hi = hi2[side];
                    c.slot = (id << 1) | side;
                    c.first_chunk = c.nchunks = 0;
                    c.depth = g.depth + 1;
                    for (int k = 0; k < 3; ++k) {
                        c.blo[k] = side ? rlo[k] : llo[k];
                        c.bhi[k] = side ? rhi[k] : lhi[k];
                    }
                    h[nh++] = c;
                }
                if (nh == 2 && (h[0].hi - h[0].lo) < (h[1].hi - h[1].lo)) {
                    const SahSeg tmp = h[0];
                    h[0] = h[1];
                    h[1] = tmp;
                }
                for (int j = 0; j < nh; ++j) {
                    if (sp + j < SAH_STACK) st[sp + j] = h[j];
                    else atomicAdd(&a.ctr[SC_ERR], 1u);
                }
            }
            int nh = (mid - g.lo > a.leaf) + (g.hi - mid > a.leaf);
            sp = min(sp + nh, SAH_STACK);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (bn[wv] > SAH_BBUF - 2) flush();  // room for the next split's two halves
        }
    }
    flush();
}

// subtrees of 2..leaf_size points: index-median splits, one thread each
__global__ void sah_bucket_kernel(SahArgs a) {
    const unsigned nb = min(a.ctr[SC_BUCKET], (unsigned)a.cap_bucket);
    for (unsigned j = blockIdx.x * blockDim.x + threadIdx.x; j < nb; j += gridDim.x * blockDim.x) {
        const int4 b = a.buckets[j];
        int2 stk[8];
        int slots[8];
        int sp = 0;
        stk[sp] = make_int2(b.x, b.y);
        slots[sp++] = b.z;
        while (sp > 0) {
            --sp;
            const int lo = stk[sp].x, hi = stk[sp].y, slot = slots[sp];
            const int m = hi - lo;
            if (m == 1) {
                a.parent[(a.n - 1) + lo] = parent_link(slot);
                continue;
            }
            const int mid = lo + m / 2;
            const int id = sah_node(a, lo, mid, hi, slot);
            if (sp + 2 > 8) {
                atomicAdd(&a.ctr[SC_ERR], 1u);
                break;
            }
            stk[sp] = make_int2(mid, hi);
            slots[sp++] = (id << 1) | 1;
            stk[sp] = make_int2(lo, mid);
            slots[sp++] = id << 1;
        }
    }
}

// ---------------------------------------------------------------------------
// host driver: after the Morton sort and gather of tree_build, instead of the
// Karras pass; the refit that follows fills the child boxes
// ---------------------------------------------------------------------------
int sah_build(Ctx* ctx, Tree* t) {
    const int n = t->n, nclouds = t->nclouds;
    if (n < 2) return GPRT_OK;
    hipStream_t s = ctx->stream;
    SahArgs a{};
    a.n = n;
    a.leaf = t->leaf_size;
    a.cap_seg = n / SAH_SMALL + nclouds + 2;
    a.cap_small = n / (a.leaf + 1) + nclouds + 2;
    a.cap_bucket = n / 2 + nclouds + 2;
    a.cap_chunks = n / SAH_CH + a.cap_seg + 2;
    // the other buffers of the partitions (the Morton sort is done with these)
    GPRT_TRY(ctx->sort_tmp[0].reserve(24ull * n));
    GPRT_TRY(ctx->sort_tmp[1].reserve(4ull * n));
    const size_t seg_b = sizeof(SahSeg) * (size_t)a.cap_seg, small_b = sizeof(SahSeg) * (size_t)a.cap_small;
    const size_t bucket_b = 16ull * a.cap_bucket, chunk_b = sizeof(SahChunk) * (size_t)a.cap_chunks;
    const size_t cnt_b = 4ull * 3 * SAH_BINS * a.cap_chunks, loff_b = 4ull * a.cap_chunks;
    const size_t bins_b = 4ull * SAH_WORDS * a.cap_seg, cuts_b = sizeof(SahCut) * (size_t)a.cap_seg;
    const size_t total = 2 * seg_b + small_b + bucket_b + chunk_b + cnt_b + loff_b + bins_b + cuts_b + 256;
    GPRT_TRY(ctx->sort_tmp[3].reserve(total));
    char* p = ctx->sort_tmp[3].as<char>();
    auto take = [&](size_t b) {
        char* r = p;
        p += (b + 15) & ~(size_t)15;
        return r;
    };
    a.ctr = (unsigned*)take(64);
    a.cur = (SahSeg*)take(seg_b);
    a.next = (SahSeg*)take(seg_b);
    a.small = (SahSeg*)take(small_b);
    a.buckets = (int4*)take(bucket_b);
    a.chunks = (SahChunk*)take(chunk_b);
    a.chunk_cnt = (uint32_t*)take(cnt_b);
    a.chunk_loff = (uint32_t*)take(loff_b);
    a.bins = (uint32_t*)take(bins_b);
    a.cuts = (SahCut*)take(cuts_b);
    a.pts = t->pts4.as<float4>();
    a.perm = t->perm.as<uint32_t>();
    a.codes = t->codes.as<uint64_t>();
    a.apts = ctx->sort_tmp[0].as<float4>();
    a.acodes = (uint64_t*)(ctx->sort_tmp[0].as<char>() + 16ull * n);
    a.aperm = ctx->sort_tmp[1].as<uint32_t>();
    a.nodes = t->nodes.as<Node>();
    a.parent = t->parent.as<int>();
    a.ranges = t->ranges.as<int2>();
    a.roots = t->roots.as<int2>();

    GPRT_TRY(launch_fill_u32(s, a.ctr, SC_WORDS, 0u));
    const int cb = std::max(1, (nclouds + 255) / 256);
    hipLaunchKernelGGL(sah_unused_kernel, dim3(cb), dim3(256), 0, s, a, t->cloud_off.as<int64_t>(), nclouds);
    hipLaunchKernelGGL(sah_roots_kernel, dim3(cb), dim3(256), 0, s, a, t->cloud_off.as<int64_t>(),
                       t->cloud_bbox.as<float>(), nclouds);
    unsigned* h = ctx->hstage;  // pinned
    for (int level = 0;; ++level) {
        if (level > 4096) return fail(GPRT_ERR_STATE, "SAH build: too many levels");
        std::swap(a.cur, a.next);
        hipLaunchKernelGGL(sah_chunklist_kernel, dim3(1), dim3(1024), 0, s, a);
        GPRT_HIP_TRY(hipMemcpyAsync(h, a.ctr, 4 * SC_WORDS, hipMemcpyDeviceToHost, s));
        GPRT_HIP_TRY(hipStreamSynchronize(s));
        const unsigned ncur = h[SC_CUR], nch = h[SC_CHUNKS];
        if (h[SC_ERR]) return fail(GPRT_ERR_STATE, "SAH build: a work list overflowed");
        if (ncur == 0) break;
        hipLaunchKernelGGL(sah_bin_kernel, dim3(nch), dim3(SAH_CT), 0, s, a);
        hipLaunchKernelGGL(sah_split_kernel, dim3(ncur), dim3(64), 0, s, a);
        hipLaunchKernelGGL(sah_scatter_kernel, dim3(nch), dim3(SAH_CT), 0, s, a);
        hipLaunchKernelGGL(sah_copyback_kernel, dim3(nch), dim3(SAH_CT), 0, s, a);
    }
    const int small_grid = std::max(1, std::min((a.cap_small + 3) / 4, 4 * ctx->num_cus));
    hipLaunchKernelGGL(sah_small_kernel, dim3(small_grid), dim3(256), 0, s, a);
    const int bucket_grid = std::max(1, std::min((a.cap_bucket + 255) / 256, 8 * ctx->num_cus));
    hipLaunchKernelGGL(sah_bucket_kernel, dim3(bucket_grid), dim3(256), 0, s, a);
    GPRT_HIP_TRY(hipGetLastError());
    GPRT_HIP_TRY(hipMemcpyAsync(h, a.ctr, 4 * SC_WORDS, hipMemcpyDeviceToHost, s));
    GPRT_HIP_TRY(hipStreamSynchronize(s));
    if (h[SC_ERR]) return fail(GPRT_ERR_STATE, "SAH build: a work list or stack overflowed");
    return GPRT_OK;
}

}  // namespace gprt
