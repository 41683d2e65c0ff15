// knn.hip -- exact k-NN traversal kernels over the LBVH.
//
//   * radius samplers (LCG of gprt-optix/src/dispatch.rs:65-103, strided of
//     gprt-macros/src/lib.rs:417-450) evaluated on the device;
//   * query ordering: Morton keys + radix sort + gather into float4;
//   * knn_round_kernel: one TrueKNN round (gprt-macros/src/lib.rs:466-594) as a
//     persistent launch.  FP32 traversal with a register candidate list of
//     capacity C >= k + guard slots, FP64 re-rank of the guard band, on-device
//     compaction of unresolved queries (replaces the host filter of
//     lib.rs:517-583 and the host radius loop of :592-594);
//   * knn_exact_kernel: FP64 traversal with an exact (d64, id) list, used as the
//     plain single-launch form and as the fallback for queries whose guard band
//     overflowed in the FP32 pass.
#include <algorithm>
#include <cfloat>

#include "knn_device.hpp"
#include "radix.hpp"

namespace gprt {

// ---------------------------------------------------------------------------
// LCG with O(log n) skip-ahead (Brown, "Random number generation with
// arbitrary strides", 1994).  Constants of dispatch.rs:66-67 (Knuth's MMIX).
// ---------------------------------------------------------------------------
#define LCG_A 6364136223846793005ull
#define LCG_C 1442695040888963407ull
#define LCG_SEED 0x123456789ABCDEFull

__device__ __forceinline__ uint64_t lcg_jump(uint64_t s, uint64_t m) {
    uint64_t am = 1, ap = 0, cm = LCG_A, cp = LCG_C;
    while (m) {
        if (m & 1) {
            am *= cm;
            ap = ap * cm + cp;
        }
        cp = (cm + 1) * cp;
        cm *= cm;
        m >>= 1;
    }
    return am * s + ap;
}

// the sampler's own FP32 distance: (dx*dx + dy*dy) + dz*dz, no fma, as the
// reference's `(q.x-p.x).powi(2) + ...` (dispatch.rs:80, lib.rs:434)
__device__ __forceinline__ float d2_sampler(const float* q, const float* p) {
    float dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
    return (dx * dx + dy * dy) + dz * dz;
}

// smallest 4 of two ascending 4-lists (bitonic merge)
__device__ __forceinline__ void merge4(float* a, const float* b) {
    float m[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) m[i] = fminf(a[i], b[3 - i]);
    float t;
#define CE(i, j) if (m[i] > m[j]) { t = m[i]; m[i] = m[j]; m[j] = t; }
    CE(0, 2) CE(1, 3) CE(0, 1) CE(2, 3)
#undef CE
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = m[i];
}

__device__ __forceinline__ void ins4(float* d, float v) {
    if (v < d[3]) {
        d[3] = v;
        float t;
        if (d[3] < d[2]) { t = d[2]; d[2] = d[3]; d[3] = t; }
        if (d[2] < d[1]) { t = d[1]; d[1] = d[2]; d[2] = t; }
        if (d[1] < d[0]) { t = d[0]; d[0] = d[1]; d[1] = t; }
    }
}

// grid (100, nclouds), 1024 threads.  kth[c*100 + s] = sqrt of the 4th
// smallest sampled d^2 of sample slot s, or -1 when the reference would not
// push it.  Each lane jumps straight to its share of the draw sequence, issues
// the point loads of its (at most five) draws together, then keeps its four
// smallest values; lanes, then waves, merge their lists.
constexpr int SAMPLER_THREADS = 1024;
constexpr int SAMPLER_PER = 5;  // draws per lane: 1024 x 5 >= 5000

// The LCG advanced by SAMPLER_PER * t steps is state -> a[t] * state + c[t]
// (mod 2^64); the table is built at compile time, so lane t jumps to its
// share of a slot's draws with one multiply-add.
struct LcgJumpTable {
    uint64_t a[SAMPLER_THREADS], c[SAMPLER_THREADS];
    constexpr LcgJumpTable() : a(), c() {
        uint64_t ap = 1, cp = 0;  // one lane's stride: SAMPLER_PER steps
        for (int i = 0; i < SAMPLER_PER; ++i) {
            cp = cp * LCG_A + LCG_C;
            ap *= LCG_A;
        }
        uint64_t am = 1, cm = 0;
        for (int t = 0; t < SAMPLER_THREADS; ++t) {
            a[t] = am;
            c[t] = cm;
            cm = cm * ap + cp;
            am *= ap;
        }
    }
};
__constant__ LcgJumpTable g_lcg_jump = LcgJumpTable();

// x mod n for 0 <= x, 0 < n < 2^31, without an integer division: the FP64
// product x * fl(1/n) has relative error below 2^-51, so it is within
// 2^-21 / n of x / n; a non-integer x / n lies at least 1 / n from the
// integers around it, so the truncated product is floor(x / n), or one less
// when x / n is an integer, which the single correction step fixes.
struct ModN {
    uint32_t n;
    double r;
    __device__ explicit ModN(uint32_t n_) : n(n_), r(1.0 / (double)n_) {}
    __device__ __forceinline__ uint32_t operator()(uint32_t x) const {
        const uint32_t q = (uint32_t)((double)x * r);
        uint32_t m = x - q * n;
        if (m >= n) m -= n;
        return m;
    }
};
__global__ __launch_bounds__(SAMPLER_THREADS) void sampler_lcg_kernel(const float4* __restrict__ orig4,
                                                                      const int64_t* __restrict__ poff,
                                                                      const float* __restrict__ qs,
                                                                      const int64_t* __restrict__ qoff, float* kth) {
    const int c = blockIdx.y, s = blockIdx.x;
    const int64_t n = poff[c + 1] - poff[c], nq = qoff[c + 1] - qoff[c];
    const int64_t S = n < 5000 ? n : 5000, qss = nq < 100 ? nq : 100;
    __shared__ float part[SAMPLER_THREADS / 64][4];
    if (s >= qss || n <= 0) {
        if (threadIdx.x == 0) kth[c * 100 + s] = -1.0f;
        return;
    }
    // step index (1-based) of this slot's query draw: 1 + s*(S+1); draws are
    // < 2^31 and n, nq < 2^31, so the modulo is exact in 32 bits
    const uint64_t qstep = 1ull + (uint64_t)s * (uint64_t)(S + 1);
    const uint64_t qstate = lcg_jump(LCG_SEED, qstep);
    const int64_t qi = (int64_t)((uint32_t)(qstate >> 33) % (uint32_t)nq);
    float q[3] = {qs[3 * (qoff[c] + qi)], qs[3 * (qoff[c] + qi) + 1], qs[3 * (qoff[c] + qi) + 2]};
    constexpr int PER = SAMPLER_PER;
    const int64_t t0 = (int64_t)threadIdx.x * PER;
    float d[4] = {FLT_MAX, FLT_MAX, FLT_MAX, FLT_MAX};
    if (t0 < S) {
        uint64_t st = g_lcg_jump.a[threadIdx.x] * qstate + g_lcg_jump.c[threadIdx.x];
        const ModN mod((uint32_t)n);
        uint32_t idx[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            st = st * LCG_A + LCG_C;
            idx[u] = mod((uint32_t)(st >> 33));
        }
        float4 p4[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u)
            if (t0 + u < S) p4[u] = orig4[poff[c] + idx[u]];
#pragma unroll
        for (int u = 0; u < PER; ++u)
            if (t0 + u < S) {
                const float p[3] = {p4[u].x, p4[u].y, p4[u].z};
                const float dd = d2_sampler(q, p);
                if (dd < INFINITY) ins4(d, dd);
            }
    }
    for (int o = 32; o > 0; o >>= 1) {
        float b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) b[i] = __shfl_xor(d[i], o);
        merge4(d, b);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int i = 0; i < 4; ++i) part[w][i] = d[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int ww = 1; ww < SAMPLER_THREADS / 64; ++ww) merge4(d, part[ww]);
        // the reference pushes dists[3] when it is not f32::MAX; its entries
        // are sqrt'd, and sqrt is monotone, so the 4th sqrt is sqrt(4th d^2)
        kth[c * 100 + s] = (d[3] < FLT_MAX) ? sqrtf(d[3]) : -1.0f;
    }
}

// grid (100, nclouds), 256 threads.  Strided sampler slot s: query s*qstep
// against every pstep-th point (at most 1000), k smallest d^2 with
// 1e-10 < d^2 (< f32::MAX); kth = sqrt(k-th) or -1.
__global__ __launch_bounds__(256) void sampler_strided_kernel(const float4* __restrict__ orig4,
                                                              const int64_t* __restrict__ poff,
                                                              const float* __restrict__ qs,
                                                              const int64_t* __restrict__ qoff, int k,
                                                              float* kth) {
    const int c = blockIdx.y, s = blockIdx.x;
    const int64_t n = poff[c + 1] - poff[c], nq = qoff[c + 1] - qoff[c];
    const int64_t qstep = nq / 100 > 1 ? nq / 100 : 1, pstep = n / 1000 > 1 ? n / 1000 : 1;
    const int64_t qi = (int64_t)s * qstep;
    __shared__ float v[1024];
    if (qi >= nq) {
        if (threadIdx.x == 0) kth[c * 100 + s] = -1.0f;
        return;
    }
    const float* qp = qs + 3 * (qoff[c] + qi);
    float q[3] = {qp[0], qp[1], qp[2]};
    for (int j = threadIdx.x; j < 1024; j += 256) {
        int64_t pi = (int64_t)j * pstep;
        float dd = INFINITY;
        if (j < 1000 && pi < n) {
            const float4 p4 = orig4[poff[c] + pi];
            const float p[3] = {p4.x, p4.y, p4.z};
            dd = d2_sampler(q, p);
            if (!(dd > 1e-10f && dd < FLT_MAX)) dd = INFINITY;
        }
        v[j] = dd;
    }
    __syncthreads();
    // bitonic sort of 1024 values
    for (int size = 2; size <= 1024; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int j = threadIdx.x; j < 1024; j += 256) {
                int partner = j ^ stride;
                if (partner > j) {
                    bool up = (j & size) == 0;
                    float a = v[j], b = v[partner];
                    if ((a > b) == up) { v[j] = b; v[partner] = a; }
                }
            }
            __syncthreads();
        }
    if (threadIdx.x == 0) kth[c * 100 + s] = (k <= 1024 && v[k - 1] < INFINITY) ? sqrtf(v[k - 1]) : -1.0f;
}

// one block (128 threads) per cloud: radius from the pushed samples
// (dispatch.rs:92-103 for LCG, lib.rs:442-447 for strided).  The <= 100 values
// are ordered by rank (value, slot) so the block needs no sequential sort.
__global__ __launch_bounds__(128) void sampler_finalize_kernel(const float* kth, int nclouds, int sampler,
                                                               int heuristic, float param, float* radius) {
    __shared__ float v[100];
    __shared__ float sorted[100];
    __shared__ int len_s;
    const int c = blockIdx.x, t = threadIdx.x;
    if (t < 100) v[t] = kth[c * 100 + t];
    __syncthreads();
    // rank among the pushed values (negative = not pushed)
    if (t < 100 && v[t] >= 0.0f) {
        int r = 0;
        for (int j = 0; j < 100; ++j) {
            const float x = v[j];
            r += x >= 0.0f && (x < v[t] || (x == v[t] && j < t));
        }
        sorted[r] = v[t];
    }
    if (t == 0) {
        int n = 0;
        for (int j = 0; j < 100; ++j) n += v[j] >= 0.0f;
        len_s = n;
    }
    __syncthreads();
    if (t != 0) return;
    const int len = len_s;
    float r;
    if (sampler == GPRT_SAMPLER_LCG) {
        r = 10.0f;
        if (len > 0) {
            if (heuristic == GPRT_RADIUS_MAX) r = sorted[len - 1] * 1.1f;
            else if (heuristic == GPRT_RADIUS_PERCENTILE) {
                float a = fminf((float)len * param, (float)(len - 1));
                int idx = a > 0.0f ? (int)a : 0;
                r = sorted[idx] * 1.1f;
            } else r = param;
        }
    } else {
        r = 1.0f;
        if (heuristic == GPRT_RADIUS_FIXED) r = param;
        else if (heuristic == GPRT_RADIUS_PERCENTILE && len > 0) {
            float a = (float)len * param;
            int idx = a > 0.0f ? (int)a : 0;
            if (idx > len - 1) idx = len - 1;
            r = sorted[idx] * 1.1f;
        }
    }
    if (heuristic == GPRT_RADIUS_INFINITE) r = INFINITY;
    radius[c] = r;
}

// ---------------------------------------------------------------------------
// query ordering
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t expand10(uint32_t x) {
    x &= 0x000003ffu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
__device__ __forceinline__ uint32_t quant10(float p, float mn, float ext) {
    float t = (p - mn) / ext * 1023.0f;
    uint32_t u = !(t > 0.0f) ? 0u : (t >= 4294967296.0f ? 0xFFFFFFFFu : (uint32_t)t);
    return u < 1023u ? u : 1023u;
}

__global__ void fill_u32_kernel(unsigned* p, int n, unsigned v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// exact FP64 traversal (plain single-launch form and band-overflow fallback)
// ---------------------------------------------------------------------------
// Exact (d64, id) list: the first KMAX - k slots hold (-inf, 0) sentinels that
// never move, so the k-th entry is always the last slot.
template <int KMAX>
struct ExactList {
    double d[KMAX];
    uint32_t id[KMAX];
    __device__ __forceinline__ void init(int k) {
#pragma unroll
        for (int j = 0; j < KMAX; ++j) {
            d[j] = j < KMAX - k ? -INFINITY : INFINITY;
            id[j] = j < KMAX - k ? 0u : GPRT_ID_NONE;
        }
    }
    // insert (dd, ii) which is known to beat the last slot
    __device__ __forceinline__ void insert(double dd, uint32_t ii) {
#pragma unroll
        for (int j = KMAX - 1; j >= 1; --j) {
            const bool shift = d[j - 1] > dd || (d[j - 1] == dd && id[j - 1] > ii);
            const bool here = !shift && (d[j] > dd || (d[j] == dd && id[j] > ii));
            const double nd = shift ? d[j - 1] : (here ? dd : d[j]);
            const uint32_t ni = shift ? id[j - 1] : (here ? ii : id[j]);
            d[j] = nd;
            id[j] = ni;
        }
        if (d[0] > dd || (d[0] == dd && id[0] > ii)) { d[0] = dd; id[0] = ii; }
    }
};


// limit: a known upper bound on the k-th d64 (+inf when none); points and
// boxes beyond it cannot hold a result and are skipped
template <int KMAX>
__device__ __forceinline__ bool exact_query(const Node* __restrict__ nodes, const float4* __restrict__ pts,
                                            int2 root, float qx, float qy, float qz, int k, ExactList<KMAX>& L,
                                            uint64_t& visits, uint64_t& tests, int max_steps, double limit) {
    int stk_ref[EXACT_STACK];
    float stk_d[EXACT_STACK];
    int sp = 0;
    auto scan = [&](int f, int m) {
        for (int t0 = 0; t0 < m; t0 += 4) {
            float4 p[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (t0 + u < m) p[u] = pts[f + t0 + u];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (t0 + u < m) {
                    const double dd = d2d(qx, qy, qz, p[u].x, p[u].y, p[u].z);
                    const uint32_t ii = (uint32_t)__float_as_int(p[u].w);
                    if (dd <= limit && (dd < L.d[KMAX - 1] || (dd == L.d[KMAX - 1] && ii < L.id[KMAX - 1])))
                        L.insert(dd, ii);
                }
        }
        tests += m;
    };
    if (root.y <= 0) return true;
    int ref = root.x;
    if (ref < 0) { scan(~ref, root.y); return true; }
    int steps = 0;
    while (true) {
        if (++steps > max_steps) return false;
        const Node nd = nodes[ref];
        ++visits;
        const double wd = fmin(L.d[KMAX - 1], limit);
        const double dA = box2d(qx, qy, qz, nd.alo, nd.ahi);
        const double dB = box2d(qx, qy, qz, nd.blo, nd.bhi);
        const int rA = __float_as_int(nd.alo.w), rB = __float_as_int(nd.blo.w);
        const bool lA = rA < 0, lB = rB < 0;
        // a box whose value equals the worst can still hold a smaller id
        if (lA && lB && dB < dA) {
            if (dB <= wd) scan(~rB, __float_as_int(nd.bhi.w));
            if (dA <= wd) scan(~rA, __float_as_int(nd.ahi.w));
        } else {
            if (lA && dA <= wd) scan(~rA, __float_as_int(nd.ahi.w));
            if (lB && dB <= wd) scan(~rB, __float_as_int(nd.bhi.w));
        }
        const double wg = fmin(L.d[KMAX - 1], limit);  // the scans may have lowered it
        const bool gA = !lA && dA <= wg, gB = !lB && dB <= wg;
        if (gA && gB) {
            const bool nearA = dA <= dB;
            if (sp >= EXACT_STACK) return false;
            stk_ref[sp] = nearA ? rB : rA;
            stk_d[sp] = f64_lower_f32(nearA ? dB : dA);
            ++sp;
            ref = nearA ? rA : rB;
            continue;
        }
        if (gA) { ref = rA; continue; }
        if (gB) { ref = rB; continue; }
        bool found = false;
        const double wp = fmin(L.d[KMAX - 1], limit);
        while (sp > 0) {
            --sp;
            if ((double)stk_d[sp] <= wp) { ref = stk_ref[sp]; found = true; break; }
        }
        if (!found) break;
    }
    return true;
}

// mode 0: plain form, query i of the caller's xyz array (input order)
// mode 1: fallback, sorted position fb_list[i]
// mode 2: sorted positions base + i (a k class too large for the FP32 list)
template <int KMAX, bool COUNT>
__device__ __forceinline__ void exact_one(const ExactArgs& a, int i) {
    double limit = INFINITY;
    float qx, qy, qz;
    int g, c = 0;
    if (a.mode == 0) {
        qx = a.q_xyz[3 * (int64_t)i]; qy = a.q_xyz[3 * (int64_t)i + 1]; qz = a.q_xyz[3 * (int64_t)i + 2];
        g = i;
    } else {
        if (a.mode == 1) limit = (double)__int_as_float(a.fb_list[i].y);
        const int s = a.mode == 1 ? a.fb_list[i].x : a.base + i;
        const float4 q = a.qs[s];
        qx = q.x; qy = q.y; qz = q.z;
        g = __float_as_int(q.w);
        c = a.qcloud ? a.qcloud[s] : 0;
    }
    const int k = a.kc ? a.kc[c] : a.k;
    ExactList<KMAX> L;
    L.init(k);
    uint64_t visits = 0, tests = 0;
    if (!exact_query<KMAX>(a.nodes, a.pts, a.roots[c], qx, qy, qz, k, L, visits, tests, a.max_steps, limit))
        atomicAdd(&a.counters[CNT_ERROR], 1u);
    const int64_t ob = a.out_base ? a.out_base[c] + (int64_t)(g - a.qoff[c]) * k : (int64_t)g * k;
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
        if (j >= KMAX - k) {
            a.ids_out[ob + j - (KMAX - k)] = L.id[j];
            if (a.d2_out) a.d2_out[ob + j - (KMAX - k)] = L.d[j];
        }
    if (COUNT) {
        atomicAdd(&a.stats64[0], (unsigned long long)visits);
        atomicAdd(&a.stats64[1], (unsigned long long)tests);
    }
}

template <int KMAX, bool COUNT>
__global__ __launch_bounds__(256) void knn_exact_kernel(ExactArgs a) {
    const unsigned n = a.mode == 1 ? a.counters[CNT_FALLBACK] : (unsigned)a.nq;
    for (unsigned ui = blockIdx.x * blockDim.x + threadIdx.x; ui < n; ui += gridDim.x * blockDim.x)
        exact_one<KMAX, COUNT>(a, (int)ui);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int launch_sampler(hipStream_t s, int sampler, const float4* orig4, const int64_t* poff,
                   const float* qs,
                   const int64_t* qoff, int nclouds, int k, int heuristic, float param, float* kth,
                   float* radius) {
    if (heuristic != GPRT_RADIUS_INFINITE && !(sampler == GPRT_SAMPLER_STRIDED && heuristic != GPRT_RADIUS_PERCENTILE)) {
        if (sampler == GPRT_SAMPLER_LCG)
            hipLaunchKernelGGL(sampler_lcg_kernel, dim3(100, nclouds), dim3(SAMPLER_THREADS), 0, s, orig4, poff,
                               qs, qoff, kth);
        else
            hipLaunchKernelGGL(sampler_strided_kernel, dim3(100, nclouds), dim3(256), 0, s, orig4, poff, qs, qoff,
                               k, kth);
    } else {
        hipLaunchKernelGGL(fill_u32_kernel, dim3((100 * nclouds + 255) / 256), dim3(256), 0, s, (unsigned*)kth,
                           100 * nclouds, 0xBF800000u /* -1.0f */);
    }
    hipLaunchKernelGGL(sampler_finalize_kernel, dim3(nclouds), dim3(128), 0, s, kth, nclouds, sampler, heuristic,
                       param, radius);
    return hipGetLastError() == hipSuccess ? GPRT_OK : fail(GPRT_ERR_HIP, "sampler launch failed");
}

static int g_query_sort_bits = 16;
int query_sort_bits() { return g_query_sort_bits; }
extern "C" GPRT_HIP_API void gprt_hip_debug_set_query_sort_bits(int b) { g_query_sort_bits = b < 8 ? 8 : (b > 30 ? 30 : b); }

// ---------------------------------------------------------------------------
// Query ordering: a 32-bit key = k class (3 bits) | cloud (cb bits) | the top
// 29 - cb bits of the 30-bit Morton code; stable radix sort of the class, the
// cloud and query_sort_bits() Morton bits, then a gather into float4 (input
// index in .w).  Queries end up grouped by k class, then by cloud, then in
// Morton order.  Input validation is folded into the key pass.
// ---------------------------------------------------------------------------
__global__ void query_key32_kernel(const float* __restrict__ qs, const int64_t* __restrict__ qoff,
                                   const float* __restrict__ bbox, int use_morton, int cb,
                                   const int* __restrict__ kc, int k_uniform, uint32_t* __restrict__ keys,
                                   uint32_t* __restrict__ vals, unsigned* flag) {
    const int c = blockIdx.y;
    const int64_t b = qoff[c], e = qoff[c + 1];
    float bb[6];
    for (int a = 0; a < 6; ++a) bb[a] = bbox[6 * c + a];
    const float ex = fmaxf(bb[3] - bb[0], 1e-6f), ey = fmaxf(bb[4] - bb[1], 1e-6f), ez = fmaxf(bb[5] - bb[2], 1e-6f);
    const uint32_t cls = (uint32_t)k_class(kc ? kc[c] : k_uniform);
    const int mbits = 29 - cb;
    const uint32_t head = (cls << 29) | ((uint32_t)c << mbits);
    bool bad = false;
    for (int64_t i = b + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float x = qs[3 * i], y = qs[3 * i + 1], z = qs[3 * i + 2];
        bad |= !(fabsf(x) <= 1e18f && fabsf(y) <= 1e18f && fabsf(z) <= 1e18f);
        uint32_t m = 0;
        if (use_morton)
            m = (expand10(quant10(x, bb[0], ex)) << 2) | (expand10(quant10(y, bb[1], ey)) << 1) |
                expand10(quant10(z, bb[2], ez));
        keys[i] = head | (m >> (30 - mbits));
        vals[i] = (uint32_t)i;
    }
    if (__ballot(bad) && __lane_id() == 0) atomicOr(flag, 1u);
}

__global__ void query_gather_kernel(const float* __restrict__ qs, const uint32_t* __restrict__ perm,
                                    const uint32_t* __restrict__ keys, int cb, int nq, float4* __restrict__ out,
                                    int* __restrict__ qcloud, unsigned* counters) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nq) {
        uint32_t g = perm[i];
        out[i] = make_float4(qs[3 * g], qs[3 * g + 1], qs[3 * g + 2], __int_as_float((int)g));
        if (qcloud) qcloud[i] = (int)((keys[i] >> (29 - cb)) & ((1u << cb) - 1u));
    }
    if (i == 0) counters[CNT_ACTIVE] = (unsigned)nq;
}

// the same key as query_key32_kernel for one cloud, generated inside the
// first radix pass's histogram kernel
struct QueryKeyGen {
    static constexpr bool kGenerates = true;
    const float* qs;
    const float* bbox;
    int use_morton;
    uint32_t head;
    int mshift;  // 30 - Morton bits kept
    unsigned* flag;
    __device__ __forceinline__ uint32_t operator()(int i, bool& bad) const {
        const float x = qs[3 * (size_t)i], y = qs[3 * (size_t)i + 1], z = qs[3 * (size_t)i + 2];
        bad |= !(fabsf(x) <= 1e18f && fabsf(y) <= 1e18f && fabsf(z) <= 1e18f);
        uint32_t m = 0;
        if (use_morton) {
            const float ex = fmaxf(bbox[3] - bbox[0], 1e-6f), ey = fmaxf(bbox[4] - bbox[1], 1e-6f),
                        ez = fmaxf(bbox[5] - bbox[2], 1e-6f);
            m = (expand10(quant10(x, bbox[0], ex)) << 2) | (expand10(quant10(y, bbox[1], ey)) << 1) |
                expand10(quant10(z, bbox[2], ez));
        }
        return head | (m >> mshift);
    }
    __device__ __forceinline__ void report(bool bad) const {
        if (__ballot(bad) && __lane_id() == 0) atomicOr(flag, 1u);
    }
};

// the last radix pass writes each query's float4 (input index in .w) and cloud
// at its sorted position: the gather fused into the scatter
struct QueryGatherSink {
    const float* qs;
    float4* out;
    int* qcloud;
    int cshift;
    uint32_t cmask;
    unsigned* counters;
    int n;
    __device__ __forceinline__ void operator()(uint32_t pos, uint32_t k, uint32_t v) const {
        out[pos] = make_float4(qs[3 * (size_t)v], qs[3 * (size_t)v + 1], qs[3 * (size_t)v + 2], __int_as_float((int)v));
        if (qcloud) qcloud[pos] = (int)((k >> cshift) & cmask);
        if (pos == 0) counters[CNT_ACTIVE] = (unsigned)n;
    }
};

static int cloud_bits_of(int nclouds) {
    int cb = 0;
    while ((1 << cb) < nclouds) ++cb;
    return cb;
}

size_t query_prep_scratch_bytes(int nclouds, int use_morton) {
    (void)nclouds;
    (void)use_morton;
    return 0;
}

// the key bits the query order sorts, [begin, end), and whether the first
// pass's histogram kernel generates the keys
struct QueryOrderBits {
    int cb, mbits, qbits, begin, end;
    bool fused_keys;
};

static QueryOrderBits query_order_bits(int nclouds, int use_morton, int nq, int sort_classes) {
    QueryOrderBits b;
    b.cb = cloud_bits_of(nclouds);
    b.mbits = 29 - b.cb;
    b.qbits = use_morton ? std::min(query_sort_bits(), b.mbits) : 0;
    b.begin = b.mbits - b.qbits;
    // the class bits (29..31) only need sorting when several k classes are present
    b.end = sort_classes ? 32 : 29;
    // one cloud (so one k class, whose bits are not sorted): the keys come out
    // of the first pass's histogram kernel, without the class bits
    b.fused_keys = nclouds == 1 && !sort_classes && nq > 1 && b.begin < b.end;
    return b;
}

int query_prep_clear_range(Ctx* ctx, int nclouds, int use_morton, int nq, int sort_classes, unsigned** first,
                           int* words) {
    *first = nullptr;
    *words = 0;
    const QueryOrderBits b = query_order_bits(nclouds, use_morton, nq, sort_classes);
    if (b.cb > 20 || nq <= 1 || b.begin >= b.end) return GPRT_OK;
    return radix_sort_prefix(ctx, nq, radix_passes(b.begin, b.end), first, words);
}

int launch_query_prep(Ctx* ctx, const float* qs, const int64_t* qoff, int nclouds, const float* bbox,
                      int use_morton, int nq, uint64_t* keys, uint32_t* perm, float4* qs4, unsigned* counters,
                      int64_t max_per_cloud, int* qcloud, unsigned* scratch, unsigned* flag, const int* kc,
                      int k_uniform, int sort_classes, bool prefix_cleared) {
    (void)scratch;
    hipStream_t s = ctx->stream;
    const QueryOrderBits b = query_order_bits(nclouds, use_morton, nq, sort_classes);
    if (b.cb > 20) return fail(GPRT_ERR_ARG, "too many clouds (max 2^20)");
    const int cb = b.cb, mbits = b.mbits;
    uint32_t* k32 = (uint32_t*)keys;
    const QueryGatherSink fin{qs, qs4, qcloud, mbits, (1u << cb) - 1u, counters, nq};
    int passes = 0;
    unsigned* err = counters + CNT_SORTERR;
    if (b.fused_keys) {
        const QueryKeyGen gen{qs, bbox, use_morton, 0u, 30 - mbits, flag};
        GPRT_TRY((radix_sort_run<uint32_t, QueryKeyGen, QueryGatherSink>(
            ctx, k32, perm, nq, b.begin, b.end, gen, &fin, nullptr, nullptr, &passes, err, prefix_cleared)));
    } else {
        const int chunks = (int)std::max<int64_t>(1, std::min<int64_t>((max_per_cloud + 255) / 256, 4096));
        hipLaunchKernelGGL(query_key32_kernel, dim3(chunks, nclouds), dim3(256), 0, s, qs, qoff, bbox, use_morton, cb,
                           kc, k_uniform, k32, perm, flag);
        GPRT_TRY((radix_sort_run<uint32_t, KeyLoad<uint32_t>, QueryGatherSink>(
            ctx, k32, perm, nq, b.begin, b.end, KeyLoad<uint32_t>{k32}, &fin, nullptr, nullptr, &passes, err,
            prefix_cleared)));
    }
    if (passes == 0)  // nothing to sort: input order
        hipLaunchKernelGGL(query_gather_kernel, dim3(std::max(1, (nq + 255) / 256)), dim3(256), 0, s, qs, perm, k32,
                           cb, nq, qs4, qcloud, counters);
    return hipGetLastError() == hipSuccess ? GPRT_OK : fail(GPRT_ERR_HIP, "query prep launch failed");
}

int launch_fill_u32(hipStream_t s, unsigned* p, int n, unsigned v) {
    if (n > 0) hipLaunchKernelGGL(fill_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    return GPRT_OK;
}

int lane_spill_depth(int stack_depth, int cls, int ring) {
    return std::max(0, stack_depth - effective_ring(cls, ring));
}

int lane_round_grid(int cls, int ring, int grid_req, int trav) {
    return trav == GPRT_TRAV_WIDE ? wide_grid(cls, ring, grid_req) : lane_grid(cls, ring, grid_req);
}

int launch_round(hipStream_t s, const RoundArgs& a, int cls, int grid, int trav, int ring, bool count) {
    if (trav == GPRT_TRAV_WAVE) launch_packet_round(s, a, cls, grid, count);
    else if (trav == GPRT_TRAV_WIDE) launch_wide_round(s, a, cls, ring, grid, count);
    else launch_lane_round(s, a, cls, ring, grid, count);
    return hipGetLastError() == hipSuccess ? GPRT_OK : fail(GPRT_ERR_HIP, "round launch failed");
}

int launch_exact(hipStream_t s, const ExactArgs& a, int kmax, int grid_threads, bool count) {
    // the fallback list (mode 1) is short or empty: a small grid strides over it
    dim3 grid(a.mode == 1 ? std::min((grid_threads + 255) / 256, 256) : (grid_threads + 255) / 256);
    if (grid.x == 0) return GPRT_OK;
#define LX(KM)                                                                          \
    if (count) hipLaunchKernelGGL((knn_exact_kernel<KM, true>), grid, dim3(256), 0, s, a); \
    else hipLaunchKernelGGL((knn_exact_kernel<KM, false>), grid, dim3(256), 0, s, a);
    if (kmax <= 8) { LX(8) }
    else if (kmax <= 16) { LX(16) }
    else if (kmax <= 32) { LX(32) }
    else { LX(64) }
#undef LX
    return hipGetLastError() == hipSuccess ? GPRT_OK : fail(GPRT_ERR_HIP, "exact launch failed");
}

}  // namespace gprt
