// lbvh.hip -- GPU LBVH build: per-cloud bounds, 30-bit Morton keys, stable LSD
// radix sort staged through LDS, Karras hierarchy, atomic bottom-up refit.
//
// Replaces the reference's closed OptiX GAS build (gprt-optix/host/main.cpp:
// 200-248), its per-round refit kernel (kernels.cu:3-16, main.cpp:254-286)
// and the host-side Scene::build (gprt-core/src/lib.rs:135-210) / Morton sort
// (gprt-core/src/morton.rs:19-56).
#include <algorithm>
#include <vector>

#include "gprt_internal.hpp"
#include "knn_internal.hpp"
#include "radix.hpp"
#include "refit.hpp"

namespace gprt {

// ---------------------------------------------------------------------------
// bounds: float min/max through order-preserving int encoding + atomics
// ---------------------------------------------------------------------------
__device__ __forceinline__ int f2ord(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float ord2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

__global__ void bbox_init_kernel(int* enc, int nclouds) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nclouds) {
        for (int a = 0; a < 3; ++a) {
            enc[6 * i + a] = f2ord(3.402823466e38f);       // f32::MAX (morton.rs:25)
            enc[6 * i + 3 + a] = f2ord(-3.402823466e38f);  // f32::MIN (morton.rs:26)
        }
    }
}

// grid: (chunks, clouds).  Each block reduces its slice of its cloud in LDS
// and adds one value per bound with an atomic (six atomics per block).
__global__ __launch_bounds__(256) void bbox_reduce_kernel(const float* __restrict__ pts,
                                                          const int64_t* __restrict__ off, int* enc) {
    __shared__ float red[6][4];
    int c = blockIdx.y;
    int64_t b = off[c], e = off[c + 1];
    float mn[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float mx[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (int64_t i = b + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e;
         i += (int64_t)gridDim.x * blockDim.x) {
        for (int a = 0; a < 3; ++a) {
            float v = pts[3 * i + a];
            mn[a] = fminf(mn[a], v);
            mx[a] = fmaxf(mx[a], v);
        }
    }
    for (int o = 32; o > 0; o >>= 1)
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], o));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], o));
        }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; ++a) {
            red[a][w] = mn[a];
            red[3 + a][w] = mx[a];
        }
    __syncthreads();
    if (threadIdx.x < 6 && b < e) {
        const int a = threadIdx.x;
        float v = red[a][0];
        for (int ww = 1; ww < 4; ++ww) v = a < 3 ? fminf(v, red[a][ww]) : fmaxf(v, red[a][ww]);
        if (a < 3) atomicMin(&enc[6 * c + a], f2ord(v));
        else atomicMax(&enc[6 * c + a], f2ord(v));
    }
}

// Box used to normalise the sort keys: a cube of the cloud's largest extent
// rather than its own box (the per-axis normalisation of morton.rs:37-47), so
// a flat cloud (a LiDAR scene is ~240 x 240 x 6 m) is not cut along its thin
// axis as often as along its wide ones.  Keys only order the points: the
// search results do not depend on this choice.
__global__ void key_box_kernel(const float* __restrict__ bbox, int nclouds, float* __restrict__ kb) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nclouds) return;
    const float* b = bbox + 6 * c;
    float e = fmaxf(b[3] - b[0], fmaxf(b[4] - b[1], b[5] - b[2]));
    for (int a = 0; a < 3; ++a) {
        kb[6 * c + a] = b[a];
        kb[6 * c + 3 + a] = b[a] + e;
    }
}

__global__ void bbox_decode_kernel(const int* enc, float* bbox, int nclouds) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 6 * nclouds) bbox[i] = ord2f(enc[i]);
}

// ---------------------------------------------------------------------------
// Morton keys (gprt-core/src/morton.rs:4-48)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t expand_bits(uint32_t x) {
    x &= 0x000003ffu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// Rust `f32 as u32` (saturating, NaN -> 0) followed by .min(1023)
__device__ __forceinline__ uint32_t quantize(float p, float mn, float ext) {
    float t = (p - mn) / ext * 1023.0f;
    uint32_t u = !(t > 0.0f) ? 0u : (t >= 4294967296.0f ? 0xFFFFFFFFu : (uint32_t)t);
    return u < 1023u ? u : 1023u;
}

__device__ __forceinline__ uint32_t morton_of(float x, float y, float z, const float* bb) {
    float ex = fmaxf(bb[3] - bb[0], 1e-6f);
    float ey = fmaxf(bb[4] - bb[1], 1e-6f);
    float ez = fmaxf(bb[5] - bb[2], 1e-6f);
    return (expand_bits(quantize(x, bb[0], ex)) << 2) | (expand_bits(quantize(y, bb[1], ey)) << 1) |
           expand_bits(quantize(z, bb[2], ez));
}

// 21-bit spread for the wide keys: bit i of x moves to bit 3i
__device__ __forceinline__ uint64_t expand_bits21(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x001F00000000FFFFull;
    x = (x | (x << 16)) & 0x001F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// b bits per axis: q = min(sat((p - mn) / ext * (2^b - 1)), 2^b - 1), the
// reference's quantisation (morton.rs:42-47) with 1023 generalised; at b = 10
// the key is the reference's 30-bit code.
__device__ __forceinline__ uint32_t quantize_b(float p, float mn, float ext, float scale, uint32_t qmax) {
    float t = (p - mn) / ext * scale;
    uint32_t u = !(t > 0.0f) ? 0u : (t >= 4294967296.0f ? 0xFFFFFFFFu : (uint32_t)t);
    return u < qmax ? u : qmax;
}

// grid: (chunks, clouds).  key = cloud << 3b | morton(b bits per axis);
// vals = input index.
__global__ void morton_keys_kernel(const float* __restrict__ pts, const int64_t* __restrict__ off,
                                   const float* __restrict__ bbox, int b, uint64_t* __restrict__ keys,
                                   uint32_t* __restrict__ vals) {
    int c = blockIdx.y;
    int64_t lo = off[c], hi = off[c + 1];
    const float* bb = bbox + 6 * c;
    const float mn[3] = {bb[0], bb[1], bb[2]};
    const float ex[3] = {fmaxf(bb[3] - bb[0], 1e-6f), fmaxf(bb[4] - bb[1], 1e-6f), fmaxf(bb[5] - bb[2], 1e-6f)};
    const uint32_t qmax = (1u << b) - 1u;
    const float scale = (float)qmax;
    for (int64_t i = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < hi;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t m = 0;
        for (int a = 0; a < 3; ++a) m = (m << 1) | expand_bits21(quantize_b(pts[3 * i + a], mn[a], ex[a], scale, qmax));
        keys[i] = ((uint64_t)c << (3 * b)) | m;
        vals[i] = (uint32_t)i;
    }
}

__global__ void morton_codes_kernel(const float* __restrict__ pts, int n, const float* __restrict__ bbox,
                                    uint32_t* __restrict__ codes) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float bb[6];
        for (int a = 0; a < 6; ++a) bb[a] = bbox[a];
        codes[i] = morton_of(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], bb);
    }
}

// scratch of a sort of n 32-bit keys (at most four passes)
size_t radix_sort_hist_bytes(int n) { return sizeof(uint32_t) * os_scratch_words(n, 4) + 1024; }

int radix_sort_pairs_u64(Ctx* ctx, uint64_t* keys, uint32_t* vals, int n, int begin_bit, int end_bit) {
    return radix_sort_run<uint64_t, KeyLoad<uint64_t>, KVSink<uint64_t>>(
        ctx, keys, vals, n, begin_bit, end_bit, KeyLoad<uint64_t>{keys}, nullptr, nullptr, nullptr, nullptr,
        ctx->counters.as<unsigned>() + CNT_SORTERR);
}
int radix_sort_pairs_u32(Ctx* ctx, uint32_t* keys, uint32_t* vals, int n, int begin_bit, int end_bit,
                         uint32_t** res_keys, uint32_t** res_vals) {
    return radix_sort_run<uint32_t, KeyLoad<uint32_t>, KVSink<uint32_t>>(
        ctx, keys, vals, n, begin_bit, end_bit, KeyLoad<uint32_t>{keys}, nullptr, res_keys, res_vals, nullptr,
        ctx->counters.as<unsigned>() + CNT_SORTERR);
}

// ---------------------------------------------------------------------------
// gather sorted points: pts4[i] = (xyz of input perm[i], id local to its cloud)
// ---------------------------------------------------------------------------
__global__ void gather_points_kernel(const float* __restrict__ pts, const uint32_t* __restrict__ perm,
                                     const uint64_t* __restrict__ keys, int cshift, const int64_t* __restrict__ off,
                                     float4* __restrict__ out, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        uint32_t src = perm[i];
        int c = (int)(keys[i] >> cshift);
        int id = (int)((int64_t)src - off[c]);
        out[i] = make_float4(pts[3 * src], pts[3 * src + 1], pts[3 * src + 2], __int_as_float(id));
    }
}

// ---------------------------------------------------------------------------
// Karras hierarchy over the sorted keys with the index as tie-break
// ---------------------------------------------------------------------------
__device__ __forceinline__ int delta_fn(const uint64_t* __restrict__ keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = keys[i], b = keys[j];
    if (a == b) return 64 + __clz((unsigned)(i ^ j));
    return __clzll((long long)(a ^ b));
}

__global__ void karras_kernel(const uint64_t* __restrict__ keys, int cshift, int n, int leaf_size,
                              const int64_t* __restrict__ off, Node* __restrict__ nodes, int* __restrict__ parent,
                              int2* __restrict__ roots, int2* __restrict__ ranges) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    int dr = delta_fn(keys, n, i, i + 1), dl = delta_fn(keys, n, i, i - 1);
    int d = dr > dl ? 1 : -1;
    int dmin = min(dr, dl);
    int lmax = 2;
    while (delta_fn(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta_fn(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta_fn(keys, n, i, j);
    int s = 0;
    for (int div = 2;; div <<= 1) {
        int t = (l + div - 1) / div;
        if (delta_fn(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    int gamma = i + s * d + min(d, 0);
    int lo = min(i, j), hi = max(i, j);
    ranges[i] = make_int2(lo, hi);
    int cntA = gamma - lo + 1, cntB = hi - gamma;
    // child refs: single points are always leaves; small subtrees become leaf buckets
    int refA = (lo == gamma || cntA <= leaf_size) ? ~lo : gamma;
    int refB = (hi == gamma + 1 || cntB <= leaf_size) ? ~(gamma + 1) : gamma + 1;
    Node* nd = nodes + i;
    nd->alo.w = __int_as_float(refA);
    nd->ahi.w = __int_as_float(cntA);
    nd->blo.w = __int_as_float(refB);
    nd->bhi.w = __int_as_float(cntB);
    // parents for the refit, as (node << 1 | side): internal children at
    // [0, n-1), points at [n-1, 2n-1)
    parent[(lo == gamma ? (n - 1) : 0) + gamma] = (i << 1) | 0;
    parent[(hi == gamma + 1 ? (n - 1) : 0) + gamma + 1] = (i << 1) | 1;
    if (i == 0) parent[0] = -1;
    // cloud roots: the node whose range is exactly one cloud
    int c = (int)(keys[lo] >> cshift);
    if ((int64_t)lo == off[c] && (int64_t)hi == off[c + 1] - 1 && (hi - lo + 1) > leaf_size) roots[c] = make_int2(i, hi - lo + 1);
}

// 4-wide view: entry i = the slots of binary node i's children, a leaf-bucket
// child standing for itself in its first slot (the second stays empty)
__global__ void collapse4_kernel(const Node* __restrict__ nodes, int m, Node4* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const Node nd = nodes[i];
    float lo[4][3], hi[4][3];
    int ref[4], cnt[4];
    for (int h = 0; h < 2; ++h) {
        const float4 clo = h ? nd.blo : nd.alo, chi = h ? nd.bhi : nd.ahi;
        const int r = __float_as_int(clo.w);
        if (r >= 0) {
            const Node c = nodes[r];
            const float4 l0 = c.alo, h0 = c.ahi, l1 = c.blo, h1 = c.bhi;
            lo[2 * h][0] = l0.x; lo[2 * h][1] = l0.y; lo[2 * h][2] = l0.z;
            hi[2 * h][0] = h0.x; hi[2 * h][1] = h0.y; hi[2 * h][2] = h0.z;
            ref[2 * h] = __float_as_int(l0.w);
            cnt[2 * h] = __float_as_int(h0.w);
            lo[2 * h + 1][0] = l1.x; lo[2 * h + 1][1] = l1.y; lo[2 * h + 1][2] = l1.z;
            hi[2 * h + 1][0] = h1.x; hi[2 * h + 1][1] = h1.y; hi[2 * h + 1][2] = h1.z;
            ref[2 * h + 1] = __float_as_int(l1.w);
            cnt[2 * h + 1] = __float_as_int(h1.w);
        } else {
            lo[2 * h][0] = clo.x; lo[2 * h][1] = clo.y; lo[2 * h][2] = clo.z;
            hi[2 * h][0] = chi.x; hi[2 * h][1] = chi.y; hi[2 * h][2] = chi.z;
            ref[2 * h] = r;
            cnt[2 * h] = __float_as_int(chi.w);
            // empty slot: a point box at the first child's corner, never entered (cnt 0)
            lo[2 * h + 1][0] = clo.x; lo[2 * h + 1][1] = clo.y; lo[2 * h + 1][2] = clo.z;
            hi[2 * h + 1][0] = clo.x; hi[2 * h + 1][1] = clo.y; hi[2 * h + 1][2] = clo.z;
            ref[2 * h + 1] = -1;
            cnt[2 * h + 1] = 0;
        }
    }
    Node4 o;
    o.lox = make_float4(lo[0][0], lo[1][0], lo[2][0], lo[3][0]);
    o.loy = make_float4(lo[0][1], lo[1][1], lo[2][1], lo[3][1]);
    o.loz = make_float4(lo[0][2], lo[1][2], lo[2][2], lo[3][2]);
    o.hix = make_float4(hi[0][0], hi[1][0], hi[2][0], hi[3][0]);
    o.hiy = make_float4(hi[0][1], hi[1][1], hi[2][1], hi[3][1]);
    o.hiz = make_float4(hi[0][2], hi[1][2], hi[2][2], hi[3][2]);
    o.ref = make_int4(ref[0], ref[1], ref[2], ref[3]);
    o.cnt = make_int4(cnt[0], cnt[1], cnt[2], cnt[3]);
    out[i] = o;
}

__global__ void roots_small_kernel(const int64_t* __restrict__ off, int nclouds, int leaf_size, int2* roots) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nclouds) {
        int64_t cnt = off[c + 1] - off[c];
        if (cnt <= leaf_size) roots[c] = make_int2(~(int)off[c], (int)cnt);  // empty cloud: count 0
    }
}

// the points in input order as float4 (x, y, z, 0): the radius samplers draw
// input indices, and one 16-byte load per draw serves them
__global__ void input_points_kernel(const float* __restrict__ pts, float4* __restrict__ out, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float4(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], 0.0f);
}

int launch_input_points(hipStream_t s, const float* pts, float4* out, int n) {
    if (n > 0) hipLaunchKernelGGL(input_points_kernel, dim3((n + 255) / 256), dim3(256), 0, s, pts, out, n);
    return hipGetLastError() == hipSuccess ? GPRT_OK : fail(GPRT_ERR_HIP, "input points launch failed");
}

__global__ void fill_int_kernel(int* p, int n, int v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static inline int cdiv(int64_t a, int b) { return (int)((a + b - 1) / b); }


int tree_build(Ctx* ctx, const float* d_pts, const int64_t* h_off, int nclouds, int leaf_size, Tree* t,
               int axis_bits) {
    int64_t n64 = h_off[nclouds];
    if (n64 > (int64_t)0x7FFFFFF0) return fail(GPRT_ERR_ARG, "too many points for one tree");
    int n = (int)n64;
    t->n = n;
    t->nclouds = nclouds;
    t->leaf_size = leaf_size;
    t->h_cloud_off.assign(h_off, h_off + nclouds + 1);
    int cloud_bits = 0;
    while ((1 << cloud_bits) < nclouds) ++cloud_bits;
    // bits per axis: as many as fit beside the cloud index in 64 bits (21 for
    // one or two clouds), or the reference's 10 when asked for
    const int b = axis_bits > 0 ? std::min(axis_bits, (64 - cloud_bits) / 3) : std::min(21, (64 - cloud_bits) / 3);
    if (b < 1) return fail(GPRT_ERR_ARG, "too many clouds for 64-bit keys");
    t->axis_bits = b;
    t->key_bits = 3 * b + cloud_bits;
    t->max_depth_bound = t->key_bits + 32;
    hipStream_t s = ctx->stream;

    GPRT_TRY(t->cloud_off.reserve(sizeof(int64_t) * (nclouds + 1)));
    GPRT_TRY(t->cloud_bbox.reserve(sizeof(float) * 6 * nclouds));
    GPRT_TRY(t->key_bbox.reserve(sizeof(float) * 6 * nclouds));
    GPRT_TRY(t->roots.reserve(sizeof(int2) * nclouds));
    GPRT_TRY(t->pts4.reserve(sizeof(float4) * std::max(n, 1)));
    GPRT_TRY(t->nodes.reserve(sizeof(Node) * std::max(n - 1, 1)));
    if (t->build_wide) GPRT_TRY(t->nodes4.reserve(sizeof(Node4) * std::max(n - 1, 1)));
    GPRT_TRY(t->perm.reserve(sizeof(uint32_t) * std::max(n, 1)));
    GPRT_TRY(t->orig4.reserve(sizeof(float4) * std::max(n, 1)));
    GPRT_TRY(t->codes.reserve(sizeof(uint64_t) * std::max(n, 1)));
    GPRT_TRY(t->parent.reserve(sizeof(int) * std::max(2 * n - 1, 1)));
    GPRT_TRY(t->flags.reserve(sizeof(int) * std::max(n - 1, 1)));
    GPRT_TRY(t->ranges.reserve(sizeof(int2) * std::max(n - 1, 1)));
    GPRT_TRY(t->ready.reserve(sizeof(int) * (std::max(n, 1) + 4)));
    GPRT_TRY(ctx->ws[7].reserve(sizeof(int) * 6 * nclouds));
    int* enc = ctx->ws[7].as<int>();

    GPRT_HIP_TRY(hipMemcpyAsync(t->cloud_off.ptr, h_off, sizeof(int64_t) * (nclouds + 1), hipMemcpyHostToDevice, s));
    int64_t maxc = 0;
    for (int c = 0; c < nclouds; ++c) maxc = std::max(maxc, h_off[c + 1] - h_off[c]);
    int chunks = std::max(1, std::min(cdiv(maxc, 256 * 16), 512));

    hipLaunchKernelGGL(bbox_init_kernel, dim3(cdiv(nclouds, 256)), dim3(256), 0, s, enc, nclouds);
    hipLaunchKernelGGL(bbox_reduce_kernel, dim3(chunks, nclouds), dim3(256), 0, s, d_pts,
                       t->cloud_off.as<int64_t>(), enc);
    hipLaunchKernelGGL(bbox_decode_kernel, dim3(cdiv(6 * nclouds, 256)), dim3(256), 0, s, enc,
                       t->cloud_bbox.as<float>(), nclouds);
    hipLaunchKernelGGL(key_box_kernel, dim3(cdiv(nclouds, 256)), dim3(256), 0, s, t->cloud_bbox.as<float>(),
                       nclouds, t->key_bbox.as<float>());
    hipLaunchKernelGGL(roots_small_kernel, dim3(cdiv(nclouds, 256)), dim3(256), 0, s, t->cloud_off.as<int64_t>(),
                       nclouds, leaf_size, t->roots.as<int2>());
    if (n > 0) {
        hipLaunchKernelGGL(morton_keys_kernel, dim3(std::max(1, std::min(cdiv(maxc, 256), 4096)), nclouds),
                           dim3(256), 0, s, d_pts, t->cloud_off.as<int64_t>(), t->key_bbox.as<float>(), b,
                           t->codes.as<uint64_t>(), t->perm.as<uint32_t>());
        GPRT_TRY(launch_fill_u32(s, ctx->counters.as<unsigned>() + CNT_SORTERR, 1, 0u));
        GPRT_TRY(radix_sort_pairs_u64(ctx, t->codes.as<uint64_t>(), t->perm.as<uint32_t>(), n, 0, t->key_bits));
        hipLaunchKernelGGL(gather_points_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, d_pts,
                           t->perm.as<uint32_t>(), t->codes.as<uint64_t>(), 3 * b, t->cloud_off.as<int64_t>(),
                           t->pts4.as<float4>(), n);
        hipLaunchKernelGGL(input_points_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, d_pts, t->orig4.as<float4>(),
                           n);
    }
    if (n > 1) {
        hipLaunchKernelGGL(fill_int_kernel, dim3(cdiv(n - 1, 256)), dim3(256), 0, s, t->flags.as<int>(), n - 1, 0);
        int* ready = t->ready.as<int>();
        hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, ready + n, 1, 0);
        if (t->sah)
            GPRT_TRY(sah_build(ctx, t));
        else
            hipLaunchKernelGGL(karras_kernel, dim3(cdiv(n - 1, 256)), dim3(256), 0, s, t->codes.as<uint64_t>(), 3 * b,
                               n, leaf_size, t->cloud_off.as<int64_t>(), t->nodes.as<Node>(), t->parent.as<int>(),
                               t->roots.as<int2>(), t->ranges.as<int2>());
        BoxSlots slots{t->nodes.as<Node>(), t->pts4.as<float4>()};
        launch_refit<BoxP>(s, slots, t->parent.as<int>(), t->ranges.as<int2>(), n, t->flags.as<int>(), ready,
                           (unsigned*)(ready + n));
        if (t->build_wide)
            hipLaunchKernelGGL(collapse4_kernel, dim3(cdiv(n - 1, 256)), dim3(256), 0, s, t->nodes.as<Node>(), n - 1,
                               t->nodes4.as<Node4>());
    }
    GPRT_HIP_TRY(hipGetLastError());
    GPRT_HIP_TRY(hipMemcpyAsync(t->h_bbox, t->cloud_bbox.ptr, sizeof(float) * 6, hipMemcpyDeviceToHost, s));
    unsigned sort_err = 0;
    if (n > 0) GPRT_HIP_TRY(hipMemcpyAsync(&sort_err, ctx->counters.as<unsigned>() + CNT_SORTERR, 4,
                                           hipMemcpyDeviceToHost, s));
    GPRT_HIP_TRY(hipStreamSynchronize(s));
    if (sort_err) return fail(GPRT_ERR_STATE, "tree build: radix sort look-back did not complete");
    return GPRT_OK;
}

int morton_codes_dev(Ctx* ctx, const float* d_pts, int n, uint32_t* d_codes, float* d_bbox6) {
    hipStream_t s = ctx->stream;
    GPRT_TRY(ctx->ws[7].reserve(64));
    int64_t h_off[2] = {0, n};
    int* enc = ctx->ws[7].as<int>();
    int64_t* d_off = (int64_t*)(ctx->ws[7].as<char>() + 32);
    GPRT_HIP_TRY(hipMemcpyAsync(d_off, h_off, sizeof(h_off), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(bbox_init_kernel, dim3(1), dim3(64), 0, s, enc, 1);
    if (n > 0) {
        hipLaunchKernelGGL(bbox_reduce_kernel, dim3(std::max(1, std::min(cdiv(n, 2048), 2048)), 1), dim3(256), 0, s,
                           d_pts, d_off, enc);
    }
    hipLaunchKernelGGL(bbox_decode_kernel, dim3(1), dim3(64), 0, s, enc, d_bbox6, 1);
    if (n > 0)
        hipLaunchKernelGGL(morton_codes_kernel, dim3(cdiv(n, 256)), dim3(256), 0, s, d_pts, n, d_bbox6, d_codes);
    GPRT_HIP_TRY(hipGetLastError());
    return GPRT_OK;
}

}  // namespace gprt
