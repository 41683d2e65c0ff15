// range.hip -- fixed-radius range query over the LBVH (r_nn! / range_query!,
// gprt-macros/src/lib.rs:729-750; the predicate is Sphere::intersect's
// d^2 <= r^2, gprt-core/src/lib.rs:82-93).
//
// Contract: query q returns every point with d64(q,p) <= (double)r*(double)r,
// ordered by (d64, id), as a CSR list.  The reference's version returns one
// flat array of at most 1e6 ids for all queries with no per-query split
// (lib.rs:662-668); this one keeps the split.
//
// Pipeline: Morton-ordered queries -> FP32 traversal with the guard threshold
// band(fl32(r*r)) collecting candidates, each confirmed in FP64 -> per-query
// counts -> offsets (exclusive scan) -> the same traversal writing ids and
// d64 -> per-query sort of (d64, id).
#include <algorithm>
#include <cfloat>

#include "gprt_internal.hpp"
#include "knn_internal.hpp"

namespace gprt {

// per-lane walk with a private stack of EXACT_STACK entries (never overflows:
// the radix-tree depth bound).  mode 0 counts, mode 1 writes.
template <bool WRITE>
__global__ __launch_bounds__(256) void range_kernel(const Node* __restrict__ nodes, const float4* __restrict__ pts,
                                                    int2 root, const float4* __restrict__ qs, int nq, float r_all,
                                                    const float* __restrict__ radii,
                                                    int64_t* __restrict__ counts, const int64_t* __restrict__ off,
                                                    uint32_t* __restrict__ ids, double* __restrict__ d2,
                                                    int64_t capacity, int max_steps, unsigned* err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const float4 q = qs[i];
    const int g = __float_as_int(q.w);
    const float r = radii ? radii[g] : r_all;
    const double r2 = (double)r * (double)r;
    const float tau = band(r * r);
    int64_t cnt = 0;
    int64_t pos = WRITE ? off[g] : 0, end = WRITE ? min(off[g + 1], capacity) : 0;
    // a leaf range four points at a time: the loads of a group issue together
    // (indices past the range clamped to its last point, loaded but not
    // tested), and hits keep the range's order
    auto visit = [&](int f, int m) {
        for (int t0 = 0; t0 < m; t0 += 4) {
            float4 p[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) p[u] = pts[f + min(t0 + u, m - 1)];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (t0 + u < m && d2f(q.x, q.y, q.z, p[u].x, p[u].y, p[u].z) <= tau) {
                    const double dd = d2d(q.x, q.y, q.z, p[u].x, p[u].y, p[u].z);
                    if (dd <= r2) {
                        if (WRITE) {
                            if (pos < end) {
                                ids[pos] = (uint32_t)__float_as_int(p[u].w);
                                d2[pos] = dd;
                            }
                            ++pos;
                        } else {
                            ++cnt;
                        }
                    }
                }
        }
    };
    if (root.y > 0) {
        if (root.x < 0) {
            visit(~root.x, root.y);
        } else {
            int stk[EXACT_STACK];
            int sp = 0, node = root.x, steps = 0;
            while (true) {
                if (++steps > max_steps || sp >= EXACT_STACK) { atomicAdd(err, 1u); break; }
                const Node nd = nodes[node];
                const float dA = box2f(q.x, q.y, q.z, nd.alo, nd.ahi);
                const float dB = box2f(q.x, q.y, q.z, nd.blo, nd.bhi);
                const int rA = __float_as_int(nd.alo.w), rB = __float_as_int(nd.blo.w);
                const bool inA = dA <= tau, inB = dB <= tau;
                {
                    // the two buckets of a node are adjacent ranges of the sorted
                    // points (A's then B's): in-range ones are scanned as one range
                    const bool sA = rA < 0 && inA, sB = rB < 0 && inB;
                    const int mm = (sA ? __float_as_int(nd.ahi.w) : 0) + (sB ? __float_as_int(nd.bhi.w) : 0);
                    if (mm > 0) visit(sA ? ~rA : ~rB, mm);
                }
                const bool gA = rA >= 0 && inA, gB = rB >= 0 && inB;
                if (gA && gB) {
                    stk[sp++] = rB;
                    node = rA;
                    continue;
                }
                if (gA) { node = rA; continue; }
                if (gB) { node = rB; continue; }
                if (sp == 0) break;
                node = stk[--sp];
            }
        }
    }
    if (!WRITE) counts[g] = cnt;
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
// per-query sort of (d64, id).  Segments of <= 64 are sorted by one wave with
// a register bitonic network; longer ones are queued for a block each.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool pair_gt(double da, uint32_t ia, double db, uint32_t ib) {
    return da > db || (da == db && ia > ib);
}

constexpr int SEG_LDS = 2048;  // longest segment sorted in LDS in one piece

// big_list: segments longer than 64; segments longer than SEG_LDS also get a
// range of the merge scratch (big_tmp[slot], from the running total big_total)
__global__ __launch_bounds__(256) void segsort_small_kernel(const int64_t* __restrict__ off, int nq,
                                                            uint32_t* ids, double* d2, int64_t capacity,
                                                            int* big_list, int64_t* big_tmp, unsigned* big_count,
                                                            unsigned long long* big_total) {
    const int lane = threadIdx.x & 63;
    const int seg = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (seg >= nq) return;  // whole wave leaves together
    const int64_t b = off[seg], e = min(off[seg + 1], capacity);
    const int64_t m = e - b;
    if (m <= 1) return;
    if (m > 64) {
        if (lane == 0) {
            const unsigned slot = atomicAdd(big_count, 1u);
            big_list[slot] = seg;
            big_tmp[slot] = m > SEG_LDS ? (int64_t)atomicAdd(big_total, (unsigned long long)m) : -1;
        }
        return;
    }
    double d = lane < m ? d2[b + lane] : INFINITY;
    uint32_t id = lane < m ? ids[b + lane] : 0xFFFFFFFFu;
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
    if (lane < m) {
        d2[b + lane] = d;
        ids[b + lane] = id;
    }
}

// Bitonic sort of x[0..m) (m <= SEG_LDS) by (d64, id) in LDS.  Stages of
// stride >= 8 take compare-exchange pairs from LDS, one pair per lane at a
// time.  The strides 4, 2 and 1 of every size run in registers: lane t loads
// elements [8t, 8t + 8), which hold every pair of those stages, and stores them
// back once.  That saves LDS traffic and barriers on three stages per size.
constexpr int SEG_E = 8;
__device__ __forceinline__ void seg_cx(double& da, uint32_t& ia, double& db, uint32_t& ib, bool up) {
    if (pair_gt(da, ia, db, ib) == up) {
        const double td = da; da = db; db = td;
        const uint32_t ti = ia; ia = ib; ib = ti;
    }
}
// Register stages on lane t's eight elements.  first: sizes 2, 4 and 8 in full
// (each 8-element chunk sorted in its bitonic direction); otherwise the
// strides 4, 2 and 1 of `size`.
__device__ __forceinline__ void seg_small_strides(double* sd, uint32_t* si, int P, int size, bool first) {
    const int t = threadIdx.x, base = SEG_E * t;
    if (base >= P) return;
    double d[SEG_E];
    uint32_t id[SEG_E];
#pragma unroll
    for (int e = 0; e < SEG_E; ++e) {
        d[e] = sd[base + e];
        id[e] = si[base + e];
    }
    if (first) {
#pragma unroll
        for (int sz = 2; sz <= SEG_E; sz <<= 1)
#pragma unroll
            for (int stride = sz >> 1; stride > 0; stride >>= 1)
#pragma unroll
                for (int e = 0; e < SEG_E; ++e)
                    if ((e & stride) == 0)
                        seg_cx(d[e], id[e], d[e + stride], id[e + stride], ((base + e) & sz) == 0);
    } else {
        const bool up = (base & size) == 0;  // size >= 16: one direction for the eight
#pragma unroll
        for (int stride = SEG_E / 2; stride > 0; stride >>= 1)
#pragma unroll
            for (int e = 0; e < SEG_E; ++e)
                if ((e & stride) == 0) seg_cx(d[e], id[e], d[e + stride], id[e + stride], up);
    }
#pragma unroll
    for (int e = 0; e < SEG_E; ++e) {
        sd[base + e] = d[e];
        si[base + e] = id[e];
    }
}

__device__ void lds_bitonic(double* __restrict__ xd, uint32_t* __restrict__ xi, int m, double* sd, uint32_t* si) {
    int P = SEG_E;
    while (P < m) P <<= 1;
    for (int j = threadIdx.x; j < P; j += blockDim.x) {
        sd[j] = j < m ? xd[j] : INFINITY;
        si[j] = j < m ? xi[j] : 0xFFFFFFFFu;
    }
    __syncthreads();
    // sizes 2, 4 and 8 entirely in registers
    seg_small_strides(sd, si, P, SEG_E, true);
    __syncthreads();
    for (int size = 2 * SEG_E; size <= P; size <<= 1) {
        for (int stride = size >> 1; stride >= SEG_E; stride >>= 1) {
            for (int t = threadIdx.x; t < P / 2; t += blockDim.x) {
                const int j = 2 * t - (t & (stride - 1)), p = j + stride;
                const bool up = (j & size) == 0;
                const double dj = sd[j], dp = sd[p];
                const uint32_t ij = si[j], ip = si[p];
                if (pair_gt(dj, ij, dp, ip) == up) {
                    sd[j] = dp; sd[p] = dj;
                    si[j] = ip; si[p] = ij;
                }
            }
            __syncthreads();
        }
        seg_small_strides(sd, si, P, size, false);
        __syncthreads();
    }
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
        xd[j] = sd[j];
        xi[j] = si[j];
    }
    __syncthreads();
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

// One block per segment longer than 64.  Up to SEG_LDS elements: one bitonic
// network in LDS.  Longer: SEG_LDS-element chunks are sorted in LDS, then
// runs are merged pairwise, ping-ponging with the segment's range of the
// merge scratch; each element finds its output slot as its index in its own
// run plus its rank in the partner run (binary search; (d64, id) keys are
// distinct within a segment, so ranks never collide).
__global__ __launch_bounds__(512) void segsort_big_kernel(const int64_t* __restrict__ off, const int* big_list,
                                                          const int64_t* big_tmp, const unsigned* big_count,
                                                          uint32_t* ids, double* d2, int64_t capacity,
                                                          uint32_t* tmp_ids, double* tmp_d2) {
    __shared__ double sd[SEG_LDS];
    __shared__ uint32_t si[SEG_LDS];
    if (blockIdx.x >= *big_count) return;
    const int seg = big_list[blockIdx.x];
    const int64_t b = off[seg], e = min(off[seg + 1], capacity);
    const int m = (int)(e - b);
    if (m <= SEG_LDS) {
        lds_bitonic(d2 + b, ids + b, m, sd, si);
        return;
    }
    for (int c0 = 0; c0 < m; c0 += SEG_LDS) lds_bitonic(d2 + b + c0, ids + b + c0, min(SEG_LDS, m - c0), sd, si);
    double* ad = d2 + b;
    uint32_t* ai = ids + b;
    double* bd = tmp_d2 + big_tmp[blockIdx.x];
    uint32_t* bi = tmp_ids + big_tmp[blockIdx.x];
    for (int w = SEG_LDS; w < m; w <<= 1) {
        for (int j = threadIdx.x; j < m; j += blockDim.x) {
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
    if (ad != d2 + b) {
        for (int j = threadIdx.x; j < m; j += blockDim.x) {
            d2[b + j] = ad[j];
            ids[b + j] = ai[j];
        }
    }
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
int range_run(Ctx* ctx, const Tree& T, const float* d_q, int nq, float r, int64_t* d_off, uint32_t* d_ids,
              int64_t capacity, DevBuf& counts_buf, DevBuf& d2_buf, DevBuf& big_buf, DevBuf& keys_buf,
              DevBuf& perm_buf, DevBuf& qs4_buf, DevBuf& sums_buf, int64_t* h_total, const float* d_radii) {
    hipStream_t s = ctx->stream;
    unsigned* counters = ctx->counters.as<unsigned>();
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
    if (!d_ids) {
        hipLaunchKernelGGL(range_kernel<false>, dim3(grid), dim3(256), 0, s, T.nodes.as<Node>(),
                           T.pts4.as<float4>(), root, qs4_buf.as<float4>(), nq, r, d_radii, counts_buf.as<int64_t>(),
                           nullptr, nullptr, nullptr, (int64_t)0, max_steps, counters + CNT_ERROR);
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
    GPRT_TRY(d2_buf.reserve(8ull * std::max<int64_t>(capacity, 1)));
    GPRT_TRY(big_buf.reserve(4ull * (nq + 2) + 8ull * nq + 16));  // segment list, then merge-scratch offsets
    hipLaunchKernelGGL(range_kernel<true>, dim3(grid), dim3(256), 0, s, T.nodes.as<Node>(), T.pts4.as<float4>(),
                       root, qs4_buf.as<float4>(), nq, r, d_radii, nullptr, d_off, d_ids, d2_buf.as<double>(), capacity,
                       max_steps, counters + CNT_ERROR);
    // counters: [CNT_FALLBACK] long-segment count, stats64[3] their merge-scratch total
    unsigned* big_count = counters + CNT_FALLBACK;
    unsigned long long* big_total = (unsigned long long*)(counters + CNT_TOTAL) + 3;
    int64_t* big_tmp = (int64_t*)(big_buf.as<char>() + sizeof(int) * (((size_t)nq + 1) & ~(size_t)1));
    hipLaunchKernelGGL(segsort_small_kernel, dim3((nq + 3) / 4), dim3(256), 0, s, d_off, nq, d_ids,
                       d2_buf.as<double>(), capacity, big_buf.as<int>(), big_tmp, big_count, big_total);
    struct {
        unsigned nbig;
        unsigned long long total;
    } h{0, 0};
    GPRT_HIP_TRY(hipMemcpyAsync(&h.nbig, big_count, 4, hipMemcpyDeviceToHost, s));
    GPRT_HIP_TRY(hipMemcpyAsync(&h.total, big_total, 8, hipMemcpyDeviceToHost, s));
    GPRT_HIP_TRY(hipStreamSynchronize(s));
    if (h.nbig) {
        if (h.total) {
            GPRT_TRY(ctx->sort_tmp[0].reserve(8ull * h.total));
            GPRT_TRY(ctx->sort_tmp[1].reserve(4ull * h.total));
        }
        hipLaunchKernelGGL(segsort_big_kernel, dim3(h.nbig), dim3(512), 0, s, d_off, big_buf.as<int>(), big_tmp,
                           big_count, d_ids, d2_buf.as<double>(), capacity, ctx->sort_tmp[1].as<uint32_t>(),
                           ctx->sort_tmp[0].as<double>());
    }
    GPRT_HIP_TRY(hipGetLastError());
    return GPRT_OK;
}

}  // namespace gprt
