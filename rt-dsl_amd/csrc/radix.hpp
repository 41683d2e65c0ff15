// radix.hpp -- stable LSD radix sort of (key, value) pairs with 8-bit digits,
// 2048-key tiles (256 lanes x 8 keys), shared by the tree build (64-bit Morton
// keys) and the query ordering (32-bit keys).
//
// Per pass: a tile histogram per workgroup, one workgroup turning all tile
// histograms into output offsets, and a stable scatter.  Two hooks let a
// caller fuse its own work into the first and the last pass: the first
// pass's histogram can come from a kernel that also generates the keys
// (rs_keygen_hist_kernel), and the last pass's scatter can write through a
// caller's sink instead of the key / value arrays.
#pragma once
#include "gprt_internal.hpp"

namespace gprt {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_BINS = 256;
constexpr int RS_WAVES = RS_THREADS / 64;

// hist layout: tile-major [nblocks][RS_BINS] (each tile writes 1 KB
// contiguously); rs_offsets_kernel turns it into the output offset of every
// (tile, digit) pair.
template <typename K>
__global__ __launch_bounds__(RS_THREADS) void rs_hist_kernel(const K* __restrict__ keys, int n, int shift,
                                                             uint32_t mask, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[RS_BINS];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE;
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        const int idx = base + i * RS_THREADS + threadIdx.x;
        if (idx < n) atomicAdd(&h[(uint32_t)(keys[idx] >> shift) & mask], 1u);
    }
    __syncthreads();
    hist[blockIdx.x * RS_BINS + threadIdx.x] = h[threadIdx.x];
}

// The first pass's histogram from a key generator: keys[i] = gen(i, bad),
// vals[i] = i, with the same tiling as rs_hist_kernel (workgroup `tile` of
// RS_THREADS lanes, all of which call this).  gen.report(bad) is called once
// per lane with whether any of its keys came from a bad input.
template <class KeyGen>
__device__ __forceinline__ void rs_keygen_hist_tile(const KeyGen& gen, int n, int shift, uint32_t mask,
                                                    uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                    uint32_t* __restrict__ hist, int tile) {
    __shared__ uint32_t h[RS_BINS];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = tile * RS_TILE;
    bool bad = false;
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
        const int idx = base + i * RS_THREADS + threadIdx.x;
        if (idx < n) {
            const uint32_t k = gen(idx, bad);
            keys[idx] = k;
            vals[idx] = (uint32_t)idx;
            atomicAdd(&h[(k >> shift) & mask], 1u);
        }
    }
    gen.report(bad);
    __syncthreads();
    hist[tile * RS_BINS + threadIdx.x] = h[threadIdx.x];
}

template <class KeyGen>
__global__ __launch_bounds__(RS_THREADS) void rs_keygen_hist_kernel(KeyGen gen, int n, int shift, uint32_t mask,
                                                                    uint32_t* __restrict__ keys,
                                                                    uint32_t* __restrict__ vals,
                                                                    uint32_t* __restrict__ hist) {
    rs_keygen_hist_tile(gen, n, shift, mask, keys, vals, hist, blockIdx.x);
}

// One block of 1024 threads = 4 groups x 256 digits.  Group g scans tiles
// [g*nb/4, (g+1)*nb/4) per digit; the groups and the digit totals are then
// combined so hist[b][d] = (keys with digit < d) + (keys with digit d in tiles < b).
// Up to RS_REG tiles per group are held in registers, so a group waits for one
// memory latency for its reads; beyond that the tiles are read twice, RS_UNROLL
// independent loads per round trip.
template <int RS_REG = 64>
__global__ __launch_bounds__(1024) void rs_offsets_kernel(uint32_t* hist, int nblocks) {
    __shared__ uint32_t gsum[4][RS_BINS];
    __shared__ uint32_t dtot[RS_BINS];
    const int d = threadIdx.x & 255, g = threadIdx.x >> 8;
    const int per = (nblocks + 3) / 4;
    const int b0 = min(g * per, nblocks), b1 = min(b0 + per, nblocks);
    constexpr int RS_UNROLL = 8;
    const bool in_regs = per <= RS_REG;
    uint32_t r[RS_REG];
    uint32_t run = 0;
    int b = b0;
    if (in_regs) {
#pragma unroll
        for (int u = 0; u < RS_REG; ++u) r[u] = b0 + u < b1 ? hist[(b0 + u) * RS_BINS + d] : 0u;
#pragma unroll
        for (int u = 0; u < RS_REG; ++u) run += r[u];
    } else {
        for (; b + RS_UNROLL <= b1; b += RS_UNROLL) {
            uint32_t v[RS_UNROLL];
#pragma unroll
            for (int u = 0; u < RS_UNROLL; ++u) v[u] = hist[(b + u) * RS_BINS + d];
#pragma unroll
            for (int u = 0; u < RS_UNROLL; ++u) run += v[u];
        }
        for (; b < b1; ++b) run += hist[b * RS_BINS + d];
    }
    gsum[g][d] = run;
    __syncthreads();
    if (g == 0) {
        uint32_t t = 0;
        for (int gg = 0; gg < 4; ++gg) {
            const uint32_t v = gsum[gg][d];
            gsum[gg][d] = t;
            t += v;
        }
        dtot[d] = t;
    }
    __syncthreads();
    // exclusive scan of the 256 digit totals by the first four wavefronts:
    // a shuffle scan within each, then the wavefront totals
    __shared__ uint32_t wtot[4];
    if (g == 0) {
        const uint32_t tot = dtot[d];
        uint32_t inc = tot;
        const int lane = d & 63;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t x = __shfl_up(inc, o);
            if (lane >= o) inc += x;
        }
        if (lane == 63) wtot[d >> 6] = inc;
        dtot[d] = inc - tot;
    }
    __syncthreads();
    if (g == 0) {
        uint32_t before = 0;
        for (int w = 0; w < (d >> 6); ++w) before += wtot[w];
        dtot[d] += before;
    }
    __syncthreads();
    run = dtot[d] + gsum[g][d];
    if (in_regs) {
#pragma unroll
        for (int u = 0; u < RS_REG; ++u)
            if (b0 + u < b1) {
                hist[(b0 + u) * RS_BINS + d] = run;
                run += r[u];
            }
        return;
    }
    for (b = b0; b + RS_UNROLL <= b1; b += RS_UNROLL) {
        uint32_t v[RS_UNROLL];
#pragma unroll
        for (int u = 0; u < RS_UNROLL; ++u) v[u] = hist[(b + u) * RS_BINS + d];
#pragma unroll
        for (int u = 0; u < RS_UNROLL; ++u) {
            hist[(b + u) * RS_BINS + d] = run;
            run += v[u];
        }
    }
    for (; b < b1; ++b) {
        const uint32_t v = hist[b * RS_BINS + d];
        hist[b * RS_BINS + d] = run;
        run += v;
    }
}

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

// Stable scatter of one tile: ranks within a wavefront by matching the digit
// with 8 ballots, per-wave running counts in LDS, then the tile's offsets.
template <typename K, class Sink>
__global__ __launch_bounds__(RS_THREADS) void rs_scatter_kernel(const K* __restrict__ kin,
                                                                const uint32_t* __restrict__ vin, int n,
                                                                int shift, uint32_t mask,
                                                                const uint32_t* __restrict__ hist, Sink sink) {
    __shared__ uint32_t wcnt[RS_WAVES][RS_BINS];
    __shared__ uint32_t boff[RS_BINS];
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    for (int i = t; i < RS_WAVES * RS_BINS; i += RS_THREADS) (&wcnt[0][0])[i] = 0;
    boff[t] = hist[blockIdx.x * RS_BINS + t];
    __syncthreads();

    const int base = blockIdx.x * RS_TILE + w * (64 * RS_ITEMS);
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
    {
        uint32_t s = 0;
#pragma unroll
        for (int ww = 0; ww < RS_WAVES; ++ww) {
            const uint32_t c = wcnt[ww][t];
            wcnt[ww][t] = s;
            s += c;
        }
    }
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

// Sort (keys, vals)[0, n) by bits [begin_bit, end_bit).  first_hist_ready: the
// first pass's tile histograms are already in the context's histogram buffer
// (rs_keygen_hist_kernel).  fin: when given, the last pass writes through it
// and the sorted pairs are not stored.  Otherwise, with res_keys / res_vals
// the result is left in whichever of (keys, vals) and the context's
// temporaries the last pass wrote, and these receive its location; without
// them it is copied back into keys / vals.  *passes receives the pass count.
template <typename K, class Fin>
int radix_sort_run(Ctx* ctx, K* keys, uint32_t* vals, int n, int begin_bit, int end_bit, bool first_hist_ready,
                   const Fin* fin, K** res_keys, uint32_t** res_vals, int* passes_out) {
    if (res_keys) *res_keys = keys;
    if (res_vals) *res_vals = vals;
    if (passes_out) *passes_out = 0;
    if (n <= 1 || begin_bit >= end_bit) return GPRT_OK;
    const int nblocks = (n + RS_TILE - 1) / RS_TILE;
    GPRT_TRY(ctx->sort_tmp[0].reserve(sizeof(K) * (size_t)n));
    GPRT_TRY(ctx->sort_tmp[1].reserve(sizeof(uint32_t) * (size_t)n));
    GPRT_TRY(ctx->sort_tmp[2].reserve(sizeof(uint32_t) * (size_t)RS_BINS * nblocks));
    K* k0 = keys;
    uint32_t* v0 = vals;
    K* k1 = ctx->sort_tmp[0].as<K>();
    uint32_t* v1 = ctx->sort_tmp[1].as<uint32_t>();
    uint32_t* hist = ctx->sort_tmp[2].as<uint32_t>();
    hipStream_t s = ctx->stream;
    int passes = 0;
    for (int shift = begin_bit; shift < end_bit; shift += 8, ++passes) {
        const int nb = std::min(8, end_bit - shift);
        const uint32_t mask = (1u << nb) - 1u;
        const bool last = shift + 8 >= end_bit;
        if (!(passes == 0 && first_hist_ready))
            hipLaunchKernelGGL(rs_hist_kernel<K>, dim3(nblocks), dim3(RS_THREADS), 0, s, k0, n, shift, mask, hist);
        hipLaunchKernelGGL(rs_offsets_kernel<>, dim3(1), dim3(1024), 0, s, hist, nblocks);
        if (last && fin) {
            hipLaunchKernelGGL((rs_scatter_kernel<K, Fin>), dim3(nblocks), dim3(RS_THREADS), 0, s, k0, v0, n, shift,
                               mask, hist, *fin);
        } else {
            hipLaunchKernelGGL((rs_scatter_kernel<K, KVSink<K>>), dim3(nblocks), dim3(RS_THREADS), 0, s, k0, v0, n,
                               shift, mask, hist, KVSink<K>{k1, v1});
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
