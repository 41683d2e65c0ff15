// bh_walk.hpp -- the Barnes-Hut walk of one body, shared by the GPU kernel
// (bh.hip) and its host build (host_walk.cpp), so both evaluate the same
// FP32 operations in the same order.
//
// Per internal node, per child: (com.x, com.y, com.z, mass) of the child's
// bodies, merged bottom-up with mass_merge.  Opening rule for a child of size
// s (largest box edge) whose centre of mass lies at distance d from the body:
// take it whole when s < theta * d and the body lies outside the child's box,
// otherwise open it (leaf buckets are summed body by body).  Sums are FP32
// with Neumaier compensation per component.
#pragma once
#include "gprt_internal.hpp"

namespace gprt {

// per child: (com.x, com.y, com.z, mass)
struct __align__(16) BHNode {
    float4 a, b;
};

// mass-weighted merge of two (com, mass) values; a massless pair keeps the
// midpoint of the two centres
GPRT_HD float4 mass_merge(const float4& a, const float4& b) {
    const float m = a.w + b.w;
    if (m > 0.0f) {
        const float wa = a.w / m, wb = b.w / m;
        return make_float4(__builtin_fmaf(wa, a.x, wb * b.x), __builtin_fmaf(wa, a.y, wb * b.y),
                           __builtin_fmaf(wa, a.z, wb * b.z), m);
    }
    return make_float4(0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z), m);
}

// Neumaier compensated sum
struct KSum {
    float s = 0.0f, c = 0.0f;
    GPRT_HD void add(float x) {
        const float t = s + x;
        if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    GPRT_HD float value() const { return s + c; }
};

// Plummer softening:  G m (x_j - x_i) / (|x_j - x_i|^2 + eps^2)^{3/2};
// REFLAW (gprt-macros/src/lib.rs:785, 790, 799):
//   G m (x_j - x_i) / (d2 sqrt(d2 + eps^2)) for d2 = |x_j - x_i|^2 > 1e-10
template <bool REFLAW>
struct BHAcc {
    KSum ax, ay, az;
    float x, y, z, eps2, G;
    GPRT_HD void add_mass(float mx, float my, float mz, float m) {
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

// a child is taken whole when s < theta d, i.e. s^2 < theta^2 d^2, and the
// body lies outside its box
GPRT_HD bool bh_far(float x, float y, float z, const float4& lo, const float4& hi, const float4& c, float theta2) {
    const float sx = hi.x - lo.x, sy = hi.y - lo.y, sz = hi.z - lo.z;
    const float s = fmaxf(sx, fmaxf(sy, sz));
    const float dx = c.x - x, dy = c.y - y, dz = c.z - z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const bool outside = x < lo.x || x > hi.x || y < lo.y || y > hi.y || z < lo.z || z > hi.z;
    return outside && s * s < theta2 * d2;
}

struct BHCounts {
    uint64_t visits = 0, bodies = 0, accepted = 0;
};

// The walk of one body from `root`.  Mem supplies node(i) -> Node,
// bh(i) -> BHNode, pt(i) -> float4 and mass(i) -> float over the sorted
// bodies, and the callbacks on_far(child com+mass, first, count) for an
// accepted child and on_leaf(first, count) for a bucket summed body by body
// (the host build records the decisions through them).  Returns false when
// the step or stack limit is hit.
template <int STACK, bool COUNT, bool REFLAW, class Mem, class Far, class Leaf>
GPRT_HD bool bh_walk(BHAcc<REFLAW>& A, int2 root, float theta2, int max_steps, const Mem& M, BHCounts& cnt,
                     Far&& on_far, Leaf&& on_leaf) {
    auto leaf = [&](int f, int m) {
        for (int t = 0; t < m; ++t) {
            const float4 p = M.pt(f + t);
            A.add_mass(p.x, p.y, p.z, M.mass(f + t));
        }
        if (COUNT) cnt.bodies += m;
        on_leaf(f, m);
    };
    if (root.y <= 0) return true;
    if (root.x < 0) {
        leaf(~root.x, root.y);
        return true;
    }
    int stk[STACK];
    int sp = 0, node = root.x, steps = 0;
    while (true) {
        if (++steps > max_steps || sp >= STACK) return false;
        const Node nd = M.node(node);
        const BHNode b = M.bh(node);
        if (COUNT) ++cnt.visits;
        const int rA = as_int(nd.alo.w), rB = as_int(nd.blo.w);
        const int cA = as_int(nd.ahi.w);
        int go[2];
        int ngo = 0;
        if (bh_far(A.x, A.y, A.z, nd.alo, nd.ahi, b.a, theta2)) {
            A.add_mass(b.a.x, b.a.y, b.a.z, b.a.w);
            if (COUNT) ++cnt.accepted;
            on_far(node, 0, b.a);
        } else if (rA < 0) {
            leaf(~rA, cA);
        } else {
            go[ngo++] = rA;
        }
        if (bh_far(A.x, A.y, A.z, nd.blo, nd.bhi, b.b, theta2)) {
            A.add_mass(b.b.x, b.b.y, b.b.z, b.b.w);
            if (COUNT) ++cnt.accepted;
            on_far(node, 1, b.b);
        } else if (rB < 0) {
            leaf(~rB, as_int(nd.bhi.w));
        } else {
            go[ngo++] = rB;
        }
        if (ngo == 2) {
            stk[sp++] = go[1];
            node = go[0];
            continue;
        }
        if (ngo == 1) {
            node = go[0];
            continue;
        }
        if (sp == 0) return true;
        node = stk[--sp];
    }
}

}  // namespace gprt
