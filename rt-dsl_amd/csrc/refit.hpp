// refit.hpp -- bottom-up aggregation over a Karras LBVH without device-wide
// fences on every node.
//
// The sorted points are cut into chunks of REFIT_CHUNK.  Every internal node
// whose point range lies inside one chunk is completed by that chunk's
// workgroup, with the child values and arrival counters in LDS (workgroup
// scope synchronisation only).  A walk that leaves its chunk deposits its value
// in the parent's child slot in global memory and counts the arrival; nodes
// whose two inputs came from chunks are queued.  A single workgroup then
// finishes the few nodes that span chunk boundaries (at most about
// depth x number of chunks of them), again with workgroup-scope ordering only.
//
// Payload concept (P):
//   static P leaf(const float4* pts, const float4* extra, int i)
//   static P combine(const P& a, const P& b)
//   static void store(Slots& s, int node, int side, const P& v)  -- global child slot
//   static P load(const Slots& s, int node, int side)
#pragma once
#include "gprt_internal.hpp"

namespace gprt {

constexpr int REFIT_CHUNK = 1024;
constexpr int REFIT_THREADS = 256;

template <class P, class Slots>
__global__ __launch_bounds__(REFIT_THREADS) void refit_chunk_kernel(Slots slots, const int* __restrict__ parent,
                                                                    const int2* __restrict__ ranges, int n,
                                                                    int* __restrict__ gcnt, int* __restrict__ ready,
                                                                    unsigned* ready_count) {
    __shared__ P val[REFIT_CHUNK][2];
    __shared__ int cnt[REFIT_CHUNK];
    const int c0 = blockIdx.x * REFIT_CHUNK;
    const int c1 = min(c0 + REFIT_CHUNK, n);
    for (int j = threadIdx.x; j < REFIT_CHUNK; j += REFIT_THREADS) cnt[j] = 0;
    __syncthreads();
    for (int i = c0 + threadIdx.x; i < c1; i += REFIT_THREADS) {
        P v = P::leaf(slots, i);
        int link = parent[(n - 1) + i];
        while (link >= 0) {
            const int p = link >> 1, side = link & 1;
            const int2 rg = ranges[p];
            if (rg.x >= c0 && rg.y < c1) {
                val[p - c0][side] = v;
                const int old = __hip_atomic_fetch_add(&cnt[p - c0], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (old == 0) break;
                const P a = side == 0 ? v : val[p - c0][0];
                const P b = side == 1 ? v : val[p - c0][1];
                P::store(slots, p, 0, a);
                P::store(slots, p, 1, b);
                v = P::combine(a, b);
                link = parent[p];
            } else {
                // leaving the chunk: the kernel boundary publishes this store
                P::store(slots, p, side, v);
                const int old = atomicAdd(&gcnt[p], 1);
                // queued with one returning atomic per wavefront (returning
                // atomics on one counter serialise)
                const int slot = wave_append(ready_count, old == 1);
                if (slot >= 0) ready[slot] = p;
                break;
            }
        }
    }
}

template <class P, class Slots>
__global__ __launch_bounds__(1024) void refit_top_kernel(Slots slots, const int* __restrict__ parent,
                                                         int* __restrict__ gcnt, const int* __restrict__ ready,
                                                         const unsigned* ready_count) {
    const unsigned m = *ready_count;
    for (unsigned j = threadIdx.x; j < m; j += blockDim.x) {
        int p = ready[j];
        P v = P::combine(P::load(slots, p, 0), P::load(slots, p, 1));
        int link = parent[p];
        while (link >= 0) {
            const int q = link >> 1, side = link & 1;
            P::store(slots, q, side, v);
            const int old = __hip_atomic_fetch_add(&gcnt[q], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0) break;
            v = P::combine(P::load(slots, q, 0), P::load(slots, q, 1));
            link = parent[q];
        }
    }
}

// launches both kernels; gcnt must hold max(n-1,1) zeroed ints, ready n ints,
// ready_count one zeroed unsigned
template <class P, class Slots>
inline void launch_refit(hipStream_t s, const Slots& slots, const int* parent, const int2* ranges, int n, int* gcnt,
                         int* ready, unsigned* ready_count) {
    if (n < 2) return;
    const int chunks = (n + REFIT_CHUNK - 1) / REFIT_CHUNK;
    hipLaunchKernelGGL((refit_chunk_kernel<P, Slots>), dim3(chunks), dim3(REFIT_THREADS), 0, s, slots, parent, ranges,
                       n, gcnt, ready, ready_count);
    hipLaunchKernelGGL((refit_top_kernel<P, Slots>), dim3(1), dim3(1024), 0, s, slots, parent, gcnt, ready,
                       ready_count);
}

// ---- payload: the child boxes of the LBVH nodes ----
struct BoxSlots {
    Node* nodes;
    const float4* pts;
};

struct BoxP {
    float lo[3], hi[3];
    __device__ static BoxP leaf(const BoxSlots& s, int i) {
        const float4 p = s.pts[i];
        BoxP b;
        b.lo[0] = b.hi[0] = p.x;
        b.lo[1] = b.hi[1] = p.y;
        b.lo[2] = b.hi[2] = p.z;
        return b;
    }
    __device__ static BoxP combine(const BoxP& a, const BoxP& b) {
        BoxP r;
        for (int k = 0; k < 3; ++k) {
            r.lo[k] = fminf(a.lo[k], b.lo[k]);
            r.hi[k] = fmaxf(a.hi[k], b.hi[k]);
        }
        return r;
    }
    __device__ static void store(const BoxSlots& s, int node, int side, const BoxP& v) {
        Node* nd = s.nodes + node;
        float* lo = side == 0 ? &nd->alo.x : &nd->blo.x;
        float* hi = side == 0 ? &nd->ahi.x : &nd->bhi.x;
        lo[0] = v.lo[0]; lo[1] = v.lo[1]; lo[2] = v.lo[2];
        hi[0] = v.hi[0]; hi[1] = v.hi[1]; hi[2] = v.hi[2];
    }
    __device__ static BoxP load(const BoxSlots& s, int node, int side) {
        const Node* nd = s.nodes + node;
        const float* lo = side == 0 ? &nd->alo.x : &nd->blo.x;
        const float* hi = side == 0 ? &nd->ahi.x : &nd->bhi.x;
        BoxP b;
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = lo[k];
            b.hi[k] = hi[k];
        }
        return b;
    }
};

}  // namespace gprt
