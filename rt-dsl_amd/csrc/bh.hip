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
// Opening rule for a child of size s (largest box edge) whose centre of mass is
// at distance d from x_i: take it whole when s < theta * d and x_i lies outside
// the child's box; otherwise open it (leaf buckets are summed body by body).
// Accumulation is FP32 with Neumaier compensation per component.
#include <algorithm>

#include "gprt_internal.hpp"
#include "knn_internal.hpp"
#include "refit.hpp"

namespace gprt {

// per child: (com.x, com.y, com.z, mass)
struct __align__(16) BHNode {
    float4 a, b;
};

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
        MassP r;
        r.m = a.m + b.m;
        if (r.m > 0.0f) {
            const float wa = a.m / r.m, wb = b.m / r.m;
            r.cx = __builtin_fmaf(wa, a.cx, wb * b.cx);
            r.cy = __builtin_fmaf(wa, a.cy, wb * b.cy);
            r.cz = __builtin_fmaf(wa, a.cz, wb * b.cz);
        } else {
            r.cx = 0.5f * (a.cx + b.cx);
            r.cy = 0.5f * (a.cy + b.cy);
            r.cz = 0.5f * (a.cz + b.cz);
        }
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

// Neumaier compensated sum
struct KSum {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float x) {
        const float t = s + x;
        if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ float value() const { return s + c; }
};

template <bool REFLAW>
struct BHAcc {
    KSum ax, ay, az;
    float x, y, z, eps2, G;
    __device__ __forceinline__ void add_mass(float mx, float my, float mz, float m) {
        const float dx = mx - x, dy = my - y, dz = mz - z;
        float w;
        if (REFLAW) {
            const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            if (!(d2 > 1e-10f)) return;
            w = G * m / (d2 * sqrtf(d2 + eps2));
        } else {
            const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
            const float inv = 1.0f / sqrtf(r2);
            w = G * m * inv * inv * inv;
        }
        ax.add(w * dx);
        ay.add(w * dy);
        az.add(w * dz);
    }
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
    uint64_t visits = 0, bodies = 0, accepted = 0;
    auto leaf = [&](int f, int m) {
        for (int t = 0; t < m; ++t) {
            const float4 p = pts[f + t];
            A.add_mass(p.x, p.y, p.z, mass[f + t]);
        }
        bodies += m;
    };
    // a child is taken whole when s < theta d, i.e. s^2 < theta^2 d^2, and the
    // body lies outside its box
    auto far = [&](const float4& lo, const float4& hi, const float4& c) -> bool {
        const float sx = hi.x - lo.x, sy = hi.y - lo.y, sz = hi.z - lo.z;
        const float s = fmaxf(sx, fmaxf(sy, sz));
        const float dx = c.x - A.x, dy = c.y - A.y, dz = c.z - A.z;
        const float d2 = dx * dx + dy * dy + dz * dz;
        const bool outside = A.x < lo.x || A.x > hi.x || A.y < lo.y || A.y > hi.y || A.z < lo.z || A.z > hi.z;
        return outside && s * s < theta2 * d2;
    };
    if (root.y > 0) {
        if (root.x < 0) {
            leaf(~root.x, root.y);
        } else {
            int stk[EXACT_STACK];
            int sp = 0, node = root.x, steps = 0;
            while (true) {
                if (++steps > max_steps || sp >= EXACT_STACK) { atomicAdd(err, 1u); break; }
                const Node nd = nodes[node];
                const BHNode b = bh[node];
                ++visits;
                const int rA = __float_as_int(nd.alo.w), rB = __float_as_int(nd.blo.w);
                int go[2];
                int ngo = 0;
                if (far(nd.alo, nd.ahi, b.a)) {
                    A.add_mass(b.a.x, b.a.y, b.a.z, b.a.w);
                    if (COUNT) ++accepted;
                } else if (rA < 0) {
                    leaf(~rA, __float_as_int(nd.ahi.w));
                } else {
                    go[ngo++] = rA;
                }
                if (far(nd.blo, nd.bhi, b.b)) {
                    A.add_mass(b.b.x, b.b.y, b.b.z, b.b.w);
                    if (COUNT) ++accepted;
                } else if (rB < 0) {
                    leaf(~rB, __float_as_int(nd.bhi.w));
                } else {
                    go[ngo++] = rB;
                }
                if (ngo == 2) {
                    stk[sp++] = go[1];
                    node = go[0];
                    continue;
                }
                if (ngo == 1) { node = go[0]; continue; }
                if (sp == 0) break;
                node = stk[--sp];
            }
        }
    }
    const uint32_t o = perm[i];
    acc[3 * (size_t)o] = A.ax.value();
    acc[3 * (size_t)o + 1] = A.ay.value();
    acc[3 * (size_t)o + 2] = A.az.value();
    if (COUNT) {
        atomicAdd(&stats64[0], (unsigned long long)visits);
        atomicAdd(&stats64[1], (unsigned long long)bodies);
        atomicAdd(&stats64[2], (unsigned long long)accepted);
    }
}

int bh_run(Ctx* ctx, Tree& T, const float4* d_pm, int n, float theta, float eps, float G, float* d_acc,
           DevBuf& mass_buf, bool count, bool ref_law) {
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
    hipLaunchKernelGGL((bh_force_kernel<CNT, REF>), dim3((n + 255) / 256), dim3(256), 0, s, T.nodes.as<Node>(), \
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
