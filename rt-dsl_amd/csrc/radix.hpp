// radix.hpp -- stable LSD radix sort of (key, value) pairs with 8-bit digits,
// 2048-key tiles (256 lanes x 8 keys), shared by the tree build (64-bit Morton
// keys) and the query ordering (32-bit keys).
//
// One histogram kernel for the digit totals of every pass, then one scatter
// kernel per pass (decoupled look-back for the offsets of earlier tiles).  Two
// hooks let a caller fuse its own work into the first and the last kernel: the
// histogram kernel can take a key generator that also stores the keys, and the
// last pass's scatter can write through a caller's sink instead of the key /
// value arrays.
#pragma once
#include "gprt_internal.hpp"

namespace gprt {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_BINS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;

// the default destination of a pass: the other key / value arrays
template <typename K>
struct KVSink {
    K* kout;
    uint32_t* vout;
    __device__ __forceinline__ void operator()(uint32_t pos, K k, uint32_t v) const {
        kout[pos] = k;
        vout[pos] = v;
    }
};

// ---------------------------------------------------------------------------
// One-sweep passes (Merrill and Garland's decoupled look-back, as in
// Adinets and Merrill's Onesweep radix sort): the digit totals of every pass
// come from one histogram kernel over the keys, and each pass is one scatter
// kernel whose workgroups take tiles in launch order from a counter, publish
// their tile's digit counts at once and find the counts of all earlier tiles
// by walking back over the published words, instead of a separate offsets
// kernel per pass.  A status word is a 2-bit flag (aggregate / inclusive
// prefix) and a 30-bit count, so n < 2^30.
// ---------------------------------------------------------------------------
constexpr int OS_MAXP = 8;
constexpr uint32_t OS_AGG = 1u << 30, OS_INC = 2u << 30, OS_VAL = (1u << 30) - 1;
constexpr uint32_t OS_SPIN_CAP = 1u << 22;  // look-back reads of one word before giving up (never reached)

constexpr int OS_COPIES = 8;    // digit totals kept in 8 copies (tile % 8) to spread the atomics
constexpr int OS_WINDOW = 8;    // earlier tiles' words read together per look-back step

struct OsLayout {
    uint32_t* totals;   // [OS_COPIES][passes][RS_BINS] digit totals
    uint32_t* tilectr;  // [passes] tile counters
    uint32_t* status;   // [passes][ntiles][RS_BINS]
    int ntiles;
    int passes;
};
// words cleared before a sort (totals and tile counters)
inline int os_clear_words(int passes) { return OS_COPIES * passes * RS_BINS + 64; }
// words of the one-sweep scratch for n keys and `passes` passes
inline size_t os_scratch_words(int n, int passes) {
    const size_t nt = (size_t)(n + RS_TILE - 1) / RS_TILE;
    return (size_t)os_clear_words(passes) + (size_t)passes * nt * RS_BINS;
}
inline OsLayout os_layout(uint32_t* base, int n, int passes) {
    OsLayout L;
    L.totals = base;
    L.tilectr = base + (size_t)OS_COPIES * passes * RS_BINS;
    L.status = L.tilectr + 64;
    L.ntiles = (n + RS_TILE - 1) / RS_TILE;
    L.passes = passes;
    return L;
}

// keys already in memory: the histogram kernel's key source
template <typename K>
struct KeyLoad {
    const K* keys;
    static constexpr bool kGenerates = false;
    __device__ __forceinline__ K operator()(int i, bool&) const { return keys[i]; }
    __device__ __forceinline__ void report(bool) const {}
};

// Digit totals of all passes (one LDS histogram per pass, then one atomic per
// non-empty bin), and the look-back words of this tile cleared for the
// scatters.  A generating key source also stores keys[i] and vals[i] = i.
template <typename K, class Gen>
__global__ __launch_bounds__(RS_THREADS) void os_hist_kernel(Gen gen, int n, int begin_bit, int end_bit,
                                                             K* __restrict__ keys, uint32_t* __restrict__ vals,
                                                             OsLayout L) {
    __shared__ uint32_t h[OS_MAXP][RS_BINS];
    const int t = threadIdx.x;
    const int passes = (end_bit - begin_bit + 7) / 8;
    for (int p = 0; p < passes; ++p) {
        h[p][t] = 0;
        L.status[((size_t)p * L.ntiles + blockIdx.x) * RS_BINS + t] = 0;
    }
    __syncthreads();
    const int base = blockIdx.x * RS_TILE;
    bool bad = false;
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        const int idx = base + i * RS_THREADS + t;
        if (idx < n) {
            const K k = gen(idx, bad);
            if (Gen::kGenerates) {
                keys[idx] = k;
                vals[idx] = (uint32_t)idx;
            }
            for (int p = 0; p < passes; ++p) {
                const int shift = begin_bit + 8 * p;
                const uint32_t mask = (1u << min(8, end_bit - shift)) - 1u;
                atomicAdd(&h[p][(uint32_t)(k >> shift) & mask], 1u);
            }
        }
    }
    gen.report(bad);
    __syncthreads();
    uint32_t* tot = L.totals + (size_t)(blockIdx.x % OS_COPIES) * passes * RS_BINS;
    for (int p = 0; p < passes; ++p)
        if (h[p][t]) atomicAdd(&tot[p * RS_BINS + t], h[p][t]);
}

// One stable pass: the tile's ranks as in rs_scatter_kernel, the offsets of
// earlier tiles by look-back, the digit base by a scan of the totals.
template <typename K, class Sink>
__global__ __launch_bounds__(RS_THREADS) void os_scatter_kernel(const K* __restrict__ kin,
                                                                const uint32_t* __restrict__ vin, int n, int shift,
                                                                uint32_t mask, int pass, OsLayout L, Sink sink,
                                                                unsigned* err) {
    __shared__ uint32_t wcnt[RS_WAVES][RS_BINS];
    __shared__ uint32_t boff[RS_BINS];
    __shared__ uint32_t wtot[RS_WAVES];
    __shared__ int s_tile;
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    if (t == 0) s_tile = (int)atomicAdd(&L.tilectr[pass], 1u);
    for (int i = t; i < RS_WAVES * RS_BINS; i += RS_THREADS) (&wcnt[0][0])[i] = 0;
    // digit base: exclusive scan of this pass's digit totals
    uint32_t tot = 0;
#pragma unroll
    for (int c = 0; c < OS_COPIES; ++c) tot += L.totals[((size_t)c * L.passes + pass) * RS_BINS + t];
    uint32_t inc = tot;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t x = __shfl_up(inc, o);
        if (lane >= o) inc += x;
    }
    if (lane == 63) wtot[w] = inc;
    __syncthreads();
    uint32_t dbase = inc - tot;
    for (int ww = 0; ww < w; ++ww) dbase += wtot[ww];
    const int tile = s_tile;

    const int base = tile * RS_TILE + w * (64 * RS_ITEMS);
    K k[RS_ITEMS];
    uint32_t v[RS_ITEMS];
    uint32_t rank[RS_ITEMS];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        const int idx = base + i * 64 + lane;
        const bool valid = idx < n;
        k[i] = valid ? kin[idx] : (K)0;
        v[i] = valid ? vin[idx] : 0u;
        const uint32_t d = (uint32_t)(k[i] >> shift) & mask;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool b = (d >> bit) & 1u;
            const uint64_t m = __ballot(b);
            peers &= b ? m : ~m;
        }
        const uint32_t cur = valid ? wcnt[w][d] : 0u;
        rank[i] = cur + (uint32_t)__popcll(peers & lt);
        __builtin_amdgcn_wave_barrier();
        if (valid && lane == __ffsll((unsigned long long)peers) - 1) wcnt[w][d] = cur + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    uint32_t local = 0;
#pragma unroll
    for (int ww = 0; ww < RS_WAVES; ++ww) {
        const uint32_t c = wcnt[ww][t];
        wcnt[ww][t] = local;
        local += c;
    }
    // publish this tile's count of digit t, then look back for the earlier tiles'
    uint32_t* const row = L.status + (size_t)pass * L.ntiles * RS_BINS;
    uint32_t excl = 0;
    if (tile == 0) {
        __hip_atomic_store(&row[t], OS_INC | local, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(&row[(size_t)tile * RS_BINS + t], OS_AGG | local, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        // OS_WINDOW earlier words per step, loaded together, taken in order up
        // to the first inclusive prefix (done) or the first unpublished word
        // (read again from there)
        uint32_t spins = 0;
        for (int j = tile - 1; j >= 0;) {
            uint32_t w[OS_WINDOW];
#pragma unroll
            for (int u = 0; u < OS_WINDOW; ++u)
                w[u] = u <= j ? __hip_atomic_load(&row[(size_t)(j - u) * RS_BINS + t], __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                              : OS_INC;
            bool done = false;
            int used = 0;
#pragma unroll
            for (int u = 0; u < OS_WINDOW; ++u) {
                if (done || used < u) continue;
                const uint32_t f = w[u] & ~OS_VAL;
                if (f == 0) continue;  // unpublished: stop taking words
                excl += w[u] & OS_VAL;
                used = u + 1;
                done = f == OS_INC;
            }
            if (done) break;
            j -= used;
            if (used < OS_WINDOW && ++spins > OS_SPIN_CAP) {
                atomicAdd(err, 1u);
                break;
            }
        }
        __hip_atomic_store(&row[(size_t)tile * RS_BINS + t], OS_INC | (excl + local), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    boff[t] = dbase + excl;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        const int idx = base + i * 64 + lane;
        if (idx < n) {
            const uint32_t d = (uint32_t)(k[i] >> shift) & mask;
            sink(boff[d] + wcnt[w][d] + rank[i], k[i], v[i]);
        }
    }
}

static __global__ void os_clear_kernel(uint32_t* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// Sort (keys, vals)[0, n) by bits [begin_bit, end_bit) with one-sweep passes.
// gen: the keys' source for the histogram (KeyLoad, or a generator that also
// stores keys and vals = index).  fin: when given, the last pass writes through
// it and the sorted pairs are not stored.  Otherwise, with res_keys / res_vals
// Pass count of a sort of key bits [begin_bit, end_bit).
inline int radix_passes(int begin_bit, int end_bit) {
    return begin_bit >= end_bit ? 0 : (end_bit - begin_bit + 7) / 8;
}

// The words a sort of n keys over `passes` passes zeroes before its histogram
// kernel (digit totals and tile counters), for a caller that clears them in a
// kernel it launches anyway: reserves the sort's scratch, returns the first
// word and the count.
inline int radix_sort_prefix(Ctx* ctx, int n, int passes, uint32_t** first, int* words) {
    GPRT_TRY(ctx->sort_tmp[2].reserve(sizeof(uint32_t) * os_scratch_words(n, passes)));
    *first = ctx->sort_tmp[2].as<uint32_t>();
    *words = os_clear_words(passes);
    return GPRT_OK;
}

// the result is left in whichever of (keys, vals) and the context's
// temporaries the last pass wrote, and these receive its location; without
// them it is copied back into keys / vals.  *passes receives the pass count.
// err counts look-back walks that gave up (a result to be discarded).
// prefix_cleared: the caller has zeroed the radix_sort_prefix words in a
// kernel of its own, ahead of this sort on the same stream.
template <typename K, class Gen, class Fin>
int radix_sort_run(Ctx* ctx, K* keys, uint32_t* vals, int n, int begin_bit, int end_bit, const Gen& gen,
                   const Fin* fin, K** res_keys, uint32_t** res_vals, int* passes_out, unsigned* err,
                   bool prefix_cleared = false) {
    if (res_keys) *res_keys = keys;
    if (res_vals) *res_vals = vals;
    if (passes_out) *passes_out = 0;
    if (n <= 1 || begin_bit >= end_bit) return GPRT_OK;
    const int passes = radix_passes(begin_bit, end_bit);
    if (passes > OS_MAXP || n >= (1 << 30)) return fail(GPRT_ERR_ARG, "radix sort: too many keys or bits");
    const int nblocks = (n + RS_TILE - 1) / RS_TILE;
    GPRT_TRY(ctx->sort_tmp[0].reserve(sizeof(K) * (size_t)n));
    GPRT_TRY(ctx->sort_tmp[1].reserve(sizeof(uint32_t) * (size_t)n));
    GPRT_TRY(ctx->sort_tmp[2].reserve(sizeof(uint32_t) * os_scratch_words(n, passes)));
    K* k0 = keys;
    uint32_t* v0 = vals;
    K* k1 = ctx->sort_tmp[0].as<K>();
    uint32_t* v1 = ctx->sort_tmp[1].as<uint32_t>();
    const OsLayout L = os_layout(ctx->sort_tmp[2].as<uint32_t>(), n, passes);
    hipStream_t s = ctx->stream;
    if (!prefix_cleared)
        hipLaunchKernelGGL(os_clear_kernel, dim3(1), dim3(256), 0, s, L.totals, os_clear_words(passes));
    hipLaunchKernelGGL((os_hist_kernel<K, Gen>), dim3(nblocks), dim3(RS_THREADS), 0, s, gen, n, begin_bit, end_bit,
                       keys, vals, L);
    for (int p = 0; p < passes; ++p) {
        const int shift = begin_bit + 8 * p;
        const uint32_t mask = (1u << std::min(8, end_bit - shift)) - 1u;
        if (p == passes - 1 && fin) {
            hipLaunchKernelGGL((os_scatter_kernel<K, Fin>), dim3(nblocks), dim3(RS_THREADS), 0, s, k0, v0, n, shift,
                               mask, p, L, *fin, err);
        } else {
            hipLaunchKernelGGL((os_scatter_kernel<K, KVSink<K>>), dim3(nblocks), dim3(RS_THREADS), 0, s, k0, v0, n,
                               shift, mask, p, L, KVSink<K>{k1, v1}, err);
            std::swap(k0, k1);
            std::swap(v0, v1);
        }
    }
    if (passes_out) *passes_out = passes;
    GPRT_HIP_TRY(hipGetLastError());
    if (fin) return GPRT_OK;
    if (res_keys && res_vals) {
        *res_keys = k0;
        *res_vals = v0;
    } else if (k0 != keys) {
        GPRT_HIP_TRY(hipMemcpyAsync(keys, k0, sizeof(K) * (size_t)n, hipMemcpyDeviceToDevice, s));
        GPRT_HIP_TRY(hipMemcpyAsync(vals, v0, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, s));
    }
    return GPRT_OK;
}

}  // namespace gprt
