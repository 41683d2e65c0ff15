// bh.hip -- Barnes-Hut accelerations over the LBVH of the bodies.
//
// Replaces barnes_hut! (gprt-macros/src/lib.rs:752-834), whose shader walks a
// host-built octree with autorope links (gprt-core/src/lib.rs:427-540) and, as
// shipped, opens every node because the launch passes a null traversable
// (gprt-optix/host/main.cpp:393-397).  Here the same bodies are organised by
// the GPU LBVH; each internal node additionally stores, per child, the child's
// mass and centre of mass (a second bottom-up pass with the chunked refit).
//
// Force law, Plummer softening (the workload's; G and eps as parameters):
//   a_i = sum_j G m_j (x_j - x_i) / (|x_j - x_i|^2 + eps^2)^{3/2}
// or, with GPRT_BH_REFERENCE_LAW, the reference shader's law
// (gprt-macros/src/lib.rs:785, 790, 799; there eps^2 = 1e-6):
//   a_i = sum_{j : d2 > 1e-10} G m_j (x_j - x_i) / (d2 sqrt(d2 + eps^2)),
//   d2 = |x_j - x_i|^2, applied alike to bodies and to centres of mass.
// The walk itself (opening rule, force laws, compensated sums) is in
// bh_walk.hpp, shared with the host build.
#include <algorithm>

#include "gprt_internal.hpp"
#include "knn_internal.hpp"
#include "knn_device.hpp"
#include "bh_walk.hpp"
#include "refit.hpp"

namespace gprt {

struct MassSlots {
    BHNode* bh;
    const float* mass;   // sorted order
    const float4* pts;   // sorted order
};

struct MassP {
    float m, cx, cy, cz;
    __device__ static MassP leaf(const MassSlots& s, int i) {
        const float4 p = s.pts[i];
        MassP r;
        r.m = s.mass[i];
        r.cx = p.x;
        r.cy = p.y;
        r.cz = p.z;
        return r;
    }
    __device__ static MassP combine(const MassP& a, const MassP& b) {
        const float4 v = mass_merge(make_float4(a.cx, a.cy, a.cz, a.m), make_float4(b.cx, b.cy, b.cz, b.m));
        MassP r;
        r.cx = v.x; r.cy = v.y; r.cz = v.z; r.m = v.w;
        return r;
    }
    __device__ static void store(const MassSlots& s, int node, int side, const MassP& v) {
        float4* dst = side == 0 ? &s.bh[node].a : &s.bh[node].b;
        *dst = make_float4(v.cx, v.cy, v.cz, v.m);
    }
    __device__ static MassP load(const MassSlots& s, int node, int side) {
        const float4 v = side == 0 ? s.bh[node].a : s.bh[node].b;
        MassP r;
        r.cx = v.x; r.cy = v.y; r.cz = v.z; r.m = v.w;
        return r;
    }
};

__global__ void split_pos_mass_kernel(const float4* __restrict__ pm, int n, float* __restrict__ xyz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float4 v = pm[i];
        xyz[3 * i] = v.x;
        xyz[3 * i + 1] = v.y;
        xyz[3 * i + 2] = v.z;
    }
}

__global__ void validate_mass_kernel(const float4* __restrict__ pm, int n, unsigned* flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(pm[i].w >= 0.0f && pm[i].w <= 1e30f)) atomicOr(flag, 1u);
}

__global__ void gather_mass_kernel(const float4* __restrict__ pm, const uint32_t* __restrict__ perm, int n,
                                   float* __restrict__ mass) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) mass[i] = pm[perm[i]].w;
}

struct BHMemDev {
    const Node* __restrict__ nodes_;
    const BHNode* __restrict__ bh_;
    const float4* __restrict__ pts_;
    const float* __restrict__ mass_;
    __device__ __forceinline__ Node node(int i) const { return nodes_[i]; }
    __device__ __forceinline__ BHNode bh(int i) const { return bh_[i]; }
    __device__ __forceinline__ float4 pt(int i) const { return pts_[i]; }
    __device__ __forceinline__ float mass(int i) const { return mass_[i]; }
};

template <bool COUNT, bool REFLAW>
__global__ __launch_bounds__(256) void bh_force_kernel(const Node* __restrict__ nodes, const BHNode* __restrict__ bh,
                                                       const float4* __restrict__ pts, const float* __restrict__ mass,
                                                       const uint32_t* __restrict__ perm, int n, int2 root,
                                                       float theta2, float eps, float G, float* __restrict__ acc,
                                                       int max_steps, unsigned* err,
                                                       unsigned long long* stats64) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 me = pts[i];
    BHAcc<REFLAW> A;
    A.x = me.x; A.y = me.y; A.z = me.z;
    A.eps2 = eps * eps;
    A.G = G;
    const BHMemDev M{nodes, bh, pts, mass};
    BHCounts cnt;
    if (!bh_walk<EXACT_STACK, COUNT>(A, root, theta2, max_steps, M, cnt, [](int, int, const float4&) {},
                                     [](int, int) {}))
        atomicAdd(err, 1u);
    const uint32_t o = perm[i];
    acc[3 * (size_t)o] = A.ax.value();
    acc[3 * (size_t)o + 1] = A.ay.value();
    acc[3 * (size_t)o + 2] = A.az.value();
    if (COUNT) {
        atomicAdd(&stats64[0], (unsigned long long)cnt.visits);
        atomicAdd(&stats64[1], (unsigned long long)cnt.bodies);
        atomicAdd(&stats64[2], (unsigned long long)cnt.accepted);
    }
}

// The same walk for 64 Morton-adjacent bodies at once, one body per lane and
// every body keeping its own opening decisions.  The wavefront walks the union
// of its bodies' paths with wave-uniform node loads (scalar loads through the
// constant address space): at a node, each lane whose body is still inside
// the subtree tests both children as bh_walk does (take whole, or sum the
// bucket, or open); a child is descended, with the mask of the lanes that
// opened it, when any lane opened it.  A lane outside a subtree's mask does
// nothing there, so each body sees the same additions in the same order as
// in bh_walk (A child before B, the A subtree before the B subtree), and the
// accelerations equal the per-lane walk's bit for bit.  The shared stack of
// (node, lane mask) pairs lives in registers, entry e in lane e & 63 of half
// e >> 6 (written by a lane select, read by v_readlane); a bucket's bodies are
// fetched by one vector load, lane t holding body t, and broadcast with
// v_readlane.
__device__ __forceinline__ float readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

template <bool COUNT, bool REFLAW>
__global__ __launch_bounds__(256) void bh_wave_kernel(const Node* __restrict__ nodes, const BHNode* __restrict__ bh,
                                                      const float4* __restrict__ pts, const float* __restrict__ mass,
                                                      const uint32_t* __restrict__ perm, int n, int2 root,
                                                      float theta2, float eps, float G, float* __restrict__ acc,
                                                      int max_steps, unsigned* err,
                                                      unsigned long long* stats64) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = i < n;
    const uint64_t live = __ballot(valid);
    if (live == 0) return;  // a wavefront wholly past the last body
    const float4 me = valid ? pts[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    BHAcc<REFLAW> A;
    A.x = me.x; A.y = me.y; A.z = me.z;
    A.eps2 = eps * eps;
    A.G = G;
    const GPRT_AS4 Node* cn = const_view(nodes);
    const GPRT_AS4 float4* cb = const_view((const float4*)bh);
    BHCounts cnt;
    uint64_t wave_nodes = 0, wave_bodies = 0;  // loads of the wavefront (counters)
    // the bodies [f, f + m), m <= 64, summed by the lanes with `open` set
    auto bucket = [&](int f, int m, bool open) {
        if (COUNT) wave_bodies += m;
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        float pm = 0.f;
        if (lane < m) {
            p = pts[f + lane];
            pm = mass[f + lane];
        }
        for (int t = 0; t < m; ++t) {
            const float bx = readlane_f(p.x, t), by = readlane_f(p.y, t), bz = readlane_f(p.z, t);
            const float bm = readlane_f(pm, t);
            if (open) A.add_mass(bx, by, bz, bm);
        }
        if (COUNT && open) cnt.bodies += m;
    };
    bool ok = true;
    if (root.y > 0 && root.x < 0) {
        bucket(~root.x, root.y, valid);
    } else if (root.y > 0) {
        int sref0 = 0, sref1 = 0;            // stack: node
        int smlo0 = 0, smlo1 = 0, smhi0 = 0, smhi1 = 0;  // stack: lane mask
        int sp = 0, node = uniform(root.x), steps = 0;
        uint64_t mask = live;
        while (true) {
            if (++steps > max_steps) {
                ok = false;
                break;
            }
            const Node nd = ldnode(cn + node);
            const float4 ba = ld4(cb + 2 * node), bb = ld4(cb + 2 * node + 1);
            const bool act = (mask >> lane) & 1;
            if (COUNT) ++wave_nodes;
            if (COUNT && act) ++cnt.visits;
            const int rA = uniform(as_int(nd.alo.w)), rB = uniform(as_int(nd.blo.w));
            const int cA = uniform(as_int(nd.ahi.w)), cB = uniform(as_int(nd.bhi.w));
            const bool farA = act && bh_far(A.x, A.y, A.z, nd.alo, nd.ahi, ba, theta2);
            if (farA) {
                A.add_mass(ba.x, ba.y, ba.z, ba.w);
                if (COUNT) ++cnt.accepted;
            }
            const bool openA = act && !farA;
            uint64_t mA = __ballot(openA);
            if (rA < 0) {
                if (mA) bucket(~rA, cA, openA);
                mA = 0;
            }
            const bool farB = act && bh_far(A.x, A.y, A.z, nd.blo, nd.bhi, bb, theta2);
            if (farB) {
                A.add_mass(bb.x, bb.y, bb.z, bb.w);
                if (COUNT) ++cnt.accepted;
            }
            const bool openB = act && !farB;
            uint64_t mB = __ballot(openB);
            if (rB < 0) {
                if (mB) bucket(~rB, cB, openB);
                mB = 0;
            }
            if (mA && mB) {
                if (sp >= EXACT_STACK) {
                    ok = false;
                    break;
                }
                const bool mine = lane == (sp & 63);
                const int lo = (int)(uint32_t)mB, hi = (int)(uint32_t)(mB >> 32);
                if (sp < 64) {
                    sref0 = mine ? rB : sref0;
                    smlo0 = mine ? lo : smlo0;
                    smhi0 = mine ? hi : smhi0;
                } else {
                    sref1 = mine ? rB : sref1;
                    smlo1 = mine ? lo : smlo1;
                    smhi1 = mine ? hi : smhi1;
                }
                ++sp;
                node = rA;
                mask = mA;
                continue;
            }
            if (mA) {
                node = rA;
                mask = mA;
                continue;
            }
            if (mB) {
                node = rB;
                mask = mB;
                continue;
            }
            if (sp == 0) break;
            --sp;
            const int h = sp & 63;
            node = __builtin_amdgcn_readlane(sp < 64 ? sref0 : sref1, h);
            const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane(sp < 64 ? smlo0 : smlo1, h);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane(sp < 64 ? smhi0 : smhi1, h);
            mask = ((uint64_t)hi << 32) | lo;
        }
    }
    if (COUNT && lane == 0) {
        atomicAdd(&stats64[4], (unsigned long long)wave_nodes);
        atomicAdd(&stats64[5], (unsigned long long)wave_bodies);
    }
    if (!valid) return;
    if (!ok) atomicAdd(err, 1u);
    const uint32_t o = perm[i];
    acc[3 * (size_t)o] = A.ax.value();
    acc[3 * (size_t)o + 1] = A.ay.value();
    acc[3 * (size_t)o + 2] = A.az.value();
    if (COUNT) {
        atomicAdd(&stats64[0], (unsigned long long)cnt.visits);
        atomicAdd(&stats64[1], (unsigned long long)cnt.bodies);
        atomicAdd(&stats64[2], (unsigned long long)cnt.accepted);
    }
}

int bh_run(Ctx* ctx, Tree& T, const float4* d_pm, int n, float theta, float eps, float G, float* d_acc,
           DevBuf& mass_buf, bool count, bool ref_law, bool lane_walk) {
    hipStream_t s = ctx->stream;
    unsigned* counters = ctx->counters.as<unsigned>();
    GPRT_TRY(T.bh.reserve(sizeof(BHNode) * std::max(n - 1, 1)));
    GPRT_TRY(mass_buf.reserve(4ull * std::max(n, 1)));
    GPRT_TRY(launch_fill_u32(s, counters, CNT_TOTAL + 2 * CNT_STATS64, 0u));
    hipLaunchKernelGGL(gather_mass_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_pm, T.perm.as<uint32_t>(), n,
                       mass_buf.as<float>());
    if (n > 1) {
        GPRT_TRY(launch_fill_u32(s, (unsigned*)T.flags.ptr, n - 1, 0u));
        int* ready = T.ready.as<int>();
        GPRT_TRY(launch_fill_u32(s, (unsigned*)(ready + n), 1, 0u));
        MassSlots slots{T.bh.as<BHNode>(), mass_buf.as<float>(), T.pts4.as<float4>()};
        launch_refit<MassP>(s, slots, T.parent.as<int>(), T.ranges.as<int2>(), n, T.flags.as<int>(), ready,
                            (unsigned*)(ready + n));
    }
    int2 root;
    GPRT_HIP_TRY(hipMemcpyAsync(&root, T.roots.ptr, sizeof(int2), hipMemcpyDeviceToHost, s));
    GPRT_HIP_TRY(hipStreamSynchronize(s));
    const float theta2 = theta * theta;
    GPRT_HIP_TRY(hipEventRecord(ctx->ev0, s));
#define GPRT_BH_LAUNCH(CNT, REF)                                                                               \
    hipLaunchKernelGGL((lane_walk ? bh_force_kernel<CNT, REF> : bh_wave_kernel<CNT, REF>), dim3((n + 255) / 256), \
                       dim3(256), 0, s, T.nodes.as<Node>(),                                                      \
                       T.bh.as<BHNode>(), T.pts4.as<float4>(), mass_buf.as<float>(), T.perm.as<uint32_t>(), n,   \
                       root, theta2, eps, G, d_acc, T.n + 64, counters + CNT_ERROR,                              \
                       (unsigned long long*)(counters + CNT_TOTAL))
    if (count) {
        if (ref_law) GPRT_BH_LAUNCH(true, true);
        else GPRT_BH_LAUNCH(true, false);
    } else {
        if (ref_law) GPRT_BH_LAUNCH(false, true);
        else GPRT_BH_LAUNCH(false, false);
    }
#undef GPRT_BH_LAUNCH
    GPRT_HIP_TRY(hipEventRecord(ctx->ev1, s));
    GPRT_HIP_TRY(hipGetLastError());
    return GPRT_OK;
}

int bh_prepare(Ctx* ctx, const float4* d_pm, int n, float* d_xyz, unsigned* flag) {
    hipStream_t s = ctx->stream;
    if (n <= 0) return GPRT_OK;
    hipLaunchKernelGGL(split_pos_mass_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_pm, n, d_xyz);
    hipLaunchKernelGGL(validate_mass_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_pm, n, flag);
    GPRT_HIP_TRY(hipGetLastError());
    return GPRT_OK;
}

}  // namespace gprt
