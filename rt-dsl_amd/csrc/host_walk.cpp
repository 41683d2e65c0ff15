// host_walk.cpp -- a CPU build of the per-lane FP32 k-NN walk (lane_walk.hpp)
// and of the Barnes-Hut walk (bh_walk.hpp), for tests: libgprt_host.so,
// compiled host-only (no GPU code, no GPU needed).
//
// It rebuilds the LBVH exactly as lbvh.hip lays it out (same keys, stable key
// order, Karras hierarchy with the index tie-break, leaf collapse, min/max
// boxes), then runs lane_walk and rerank_sort for each query with every read
// and write of the points, nodes, LDS ring and spill column bounds-checked.
// The list capacity C is a parameter, so capacities the GPU build no longer
// uses can be exercised too.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "bh_walk.hpp"
#include "knn_internal.hpp"
#include "lane_walk.hpp"
#include "sah.hpp"

namespace {

using namespace gprt;

// ---------------------------------------------------------------------------
// LBVH, host restatement of lbvh.hip (bbox_reduce, key_box, morton_keys with b
// bits per axis, stable sort, gather, karras, refit)
// ---------------------------------------------------------------------------
struct HostTree {
    int n = 0, leaf = 8, axis_bits = 21;
    std::vector<uint64_t> keys;
    std::vector<uint32_t> perm;
    std::vector<float4> pts4;
    std::vector<Node> nodes;
    std::vector<int2> ranges;  // sorted point range [x, y] of each internal node
    std::vector<int> parent;   // (node << 1 | side) of internal nodes, then of points; -1 at the root
    int2 root{0, 0};
};

uint32_t quantize_b(float p, float mn, float ext, float scale, uint32_t qmax) {
    const float t = (p - mn) / ext * scale;
    const uint32_t u = !(t > 0.0f) ? 0u : (t >= 4294967296.0f ? 0xFFFFFFFFu : (uint32_t)t);
    return u < qmax ? u : qmax;
}

uint64_t spread21(uint32_t v) {  // bit i of v to bit 3i
    uint64_t x = 0;
    for (int i = 0; i < 21; ++i) x |= (uint64_t)((v >> i) & 1u) << (3 * i);
    return x;
}

int delta(const std::vector<uint64_t>& k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    if (k[i] == k[j]) return 64 + __builtin_clz((unsigned)(i ^ j));
    return __builtin_clzll(k[i] ^ k[j]);
}

void child_boxes(HostTree& T);
[[noreturn]] void oob(const char* what, long long idx, long long size);

void build_tree(const float* p, int n, int leaf, HostTree& T) {
    T.n = n;
    T.leaf = leaf;
    const int b = 21;  // one cloud: 21 bits per axis
    float mn[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float mx[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], p[3 * i + a]);
            mx[a] = fmaxf(mx[a], p[3 * i + a]);
        }
    const float e = fmaxf(mx[0] - mn[0], fmaxf(mx[1] - mn[1], mx[2] - mn[2]));
    float ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = fmaxf((mn[a] + e) - mn[a], 1e-6f);
    const uint32_t qmax = (1u << b) - 1u;
    const float scale = (float)qmax;
    T.keys.resize(n);
    T.perm.resize(n);
    for (int i = 0; i < n; ++i) {
        uint64_t m = 0;
        for (int a = 0; a < 3; ++a) m = (m << 1) | spread21(quantize_b(p[3 * i + a], mn[a], ext[a], scale, qmax));
        T.keys[i] = m;
        T.perm[i] = (uint32_t)i;
    }
    std::stable_sort(T.perm.begin(), T.perm.end(), [&](uint32_t x, uint32_t y) { return T.keys[x] < T.keys[y]; });
    std::vector<uint64_t> sk(n);
    T.pts4.resize(std::max(n, 1));
    for (int i = 0; i < n; ++i) {
        const uint32_t s = T.perm[i];
        sk[i] = T.keys[s];
        T.pts4[i] = make_float4(p[3 * s], p[3 * s + 1], p[3 * s + 2], as_float((int)s));
    }
    T.keys = sk;
    // a cloud of at most `leaf` points is one bucket (its nodes are still built)
    T.root = n > leaf ? make_int2(0, n) : make_int2(~0, n);
    if (n < 2) return;
    T.nodes.assign(n - 1, Node{});
    std::vector<int2> ranges(n - 1);
    std::vector<int> parent(2 * n - 1, -1);
    for (int i = 0; i < n - 1; ++i) {
        const int dr = delta(T.keys, n, i, i + 1), dl = delta(T.keys, n, i, i - 1);
        const int d = dr > dl ? 1 : -1;
        const int dmin = std::min(dr, dl);
        int lmax = 2;
        while (delta(T.keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (delta(T.keys, n, i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = delta(T.keys, n, i, j);
        int s = 0;
        for (int div = 2;; div <<= 1) {
            const int t = (l + div - 1) / div;
            if (delta(T.keys, n, i, i + (s + t) * d) > dnode) s += t;
            if (t <= 1) break;
        }
        const int gamma = i + s * d + std::min(d, 0);
        const int lo = std::min(i, j), hi = std::max(i, j);
        ranges[i] = make_int2(lo, hi);
        const int cA = gamma - lo + 1, cB = hi - gamma;
        const int rA = (lo == gamma || cA <= leaf) ? ~lo : gamma;
        const int rB = (hi == gamma + 1 || cB <= leaf) ? ~(gamma + 1) : gamma + 1;
        Node& nd = T.nodes[i];
        nd.alo.w = as_float(rA);
        nd.ahi.w = as_float(cA);
        nd.blo.w = as_float(rB);
        nd.bhi.w = as_float(cB);
        parent[(lo == gamma ? (n - 1) : 0) + gamma] = (i << 1) | 0;
        parent[(hi == gamma + 1 ? (n - 1) : 0) + gamma + 1] = (i << 1) | 1;
    }
    T.ranges = ranges;
    T.parent = parent;
    child_boxes(T);
}

// every child box of every internal node is the min/max of the points of its
// range (what the refit computes)
void child_boxes(HostTree& T) {
    const int n = T.n;
    for (int i = 0; i < n - 1; ++i) {
        if (T.ranges[i].y < T.ranges[i].x) continue;  // unused id
        Node& nd = T.nodes[i];
        const int lo = T.ranges[i].x, hi = T.ranges[i].y, cA = as_int(nd.ahi.w);
        for (int side = 0; side < 2; ++side) {
            const int f = side ? lo + cA : lo, l = side ? hi : lo + cA - 1;
            float4 bl = make_float4(3.402823466e38f, 3.402823466e38f, 3.402823466e38f, 0.f);
            float4 bh = make_float4(-3.402823466e38f, -3.402823466e38f, -3.402823466e38f, 0.f);
            for (int t = f; t <= l; ++t) {
                const float4 q = T.pts4[t];
                bl.x = fminf(bl.x, q.x); bl.y = fminf(bl.y, q.y); bl.z = fminf(bl.z, q.z);
                bh.x = fmaxf(bh.x, q.x); bh.y = fmaxf(bh.y, q.y); bh.z = fmaxf(bh.z, q.z);
            }
            float4& L = side ? nd.blo : nd.alo;
            float4& H = side ? nd.bhi : nd.ahi;
            L.x = bl.x; L.y = bl.y; L.z = bl.z;
            H.x = bh.x; H.y = bh.y; H.z = bh.z;
        }
    }
}

// ---------------------------------------------------------------------------
// SAH tree, host restatement of sah.hip: the LBVH order of build_tree, then
// segments split by sah_choose (sah.hpp) with stable partitions, down to single
// points (subtrees of at most `leaf` points, and segments at depth
// SAH_MAX_DEPTH or more, split at their index median; the walk scans buckets);
// node ids in the Karras numbering (sah_seg_id).
// ---------------------------------------------------------------------------
void build_tree_sah(const float* p, int n, int leaf, HostTree& T) {
    build_tree(p, n, leaf, T);
    if (n < 2) return;
    T.nodes.assign(n - 1, Node{});
    T.ranges.assign(n - 1, make_int2(0, -1));
    T.parent.assign(2 * n - 1, -1);
    struct Seg { int lo, hi, plink, depth; float blo[3], bhi[3]; };
    std::vector<Seg> stack;
    Seg root{0, n, -2, 0, {3.402823466e38f, 3.402823466e38f, 3.402823466e38f},
             {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f}};
    for (int i = 0; i < n; ++i) {
        const float v[3] = {T.pts4[i].x, T.pts4[i].y, T.pts4[i].z};
        for (int a = 0; a < 3; ++a) {
            root.blo[a] = sah_dec(std::min(sah_enc(root.blo[a]), sah_enc(v[a])));
            root.bhi[a] = sah_dec(std::max(sah_enc(root.bhi[a]), sah_enc(v[a])));
        }
    }
    T.root = n > leaf ? make_int2(0, n) : make_int2(~0, n);
    stack.push_back(root);
    std::vector<uint32_t> w(SAH_WORDS);
    std::vector<float4> tp;
    std::vector<uint32_t> tperm;
    std::vector<uint64_t> tkey;
    while (!stack.empty()) {
        const Seg g = stack.back();
        stack.pop_back();
        const int m = g.hi - g.lo;
        const int link = g.plink >= 0 ? g.plink : -1;
        if (m < 2) {
            if (m == 1) T.parent[(n - 1) + g.lo] = link;
            continue;
        }
        int mid = g.lo + m / 2;
        float llo[3], lhi[3], rlo[3], rhi[3];
        for (int a = 0; a < 3; ++a) { llo[a] = rlo[a] = g.blo[a]; lhi[a] = rhi[a] = g.bhi[a]; }
        if (m > leaf && g.depth < SAH_MAX_DEPTH) {
            float sc[3];
            for (int a = 0; a < 3; ++a) sc[a] = sah_scale(g.blo[a], g.bhi[a]);
            for (int j = 0; j < SAH_WORDS; ++j) w[j] = sah_init_word(j);
            for (int i = g.lo; i < g.hi; ++i) {
                const float v[3] = {T.pts4[i].x, T.pts4[i].y, T.pts4[i].z};
                for (int a = 0; a < 3; ++a) {
                    const int b = sah_bin(v[a], g.blo[a], sc[a]);
                    ++w[sah_cnt_word(a, b)];
                    for (int k = 0; k < 3; ++k) {
                        uint32_t& l = w[sah_lo_word(a, b, k)];
                        uint32_t& h = w[sah_hi_word(a, b, k)];
                        l = std::min(l, sah_enc(v[k]));
                        h = std::max(h, sah_enc(v[k]));
                    }
                }
            }
            SahSplit sp;
            if (sah_choose(w.data(), sp)) {
                tp.clear(); tperm.clear(); tkey.clear();
                for (int pass = 0; pass < 2; ++pass)
                    for (int i = g.lo; i < g.hi; ++i) {
                        const float4 q = T.pts4[i];
                        const float c = sp.axis == 0 ? q.x : (sp.axis == 1 ? q.y : q.z);
                        const bool left = sah_bin(c, g.blo[sp.axis], sc[sp.axis]) <= sp.bin;
                        if (left == (pass == 0)) {
                            tp.push_back(q);
                            tperm.push_back(T.perm[i]);
                            tkey.push_back(T.keys[i]);
                        }
                    }
                for (int i = 0; i < m; ++i) {
                    T.pts4[g.lo + i] = tp[i];
                    T.perm[g.lo + i] = tperm[i];
                    T.keys[g.lo + i] = tkey[i];
                }
                mid = g.lo + (int)sp.nleft;
                for (int a = 0; a < 3; ++a) {
                    llo[a] = sp.llo[a]; lhi[a] = sp.lhi[a];
                    rlo[a] = sp.rlo[a]; rhi[a] = sp.rhi[a];
                }
            }
        }
        const int id = sah_seg_id(g.lo, g.hi, g.plink);
        if (id < 0 || id >= n - 1) oob("sah node", id, n - 1);
        T.ranges[id] = make_int2(g.lo, g.hi - 1);
        T.parent[id] = link;
        Node& nd = T.nodes[id];
        nd.ahi.w = as_float(mid - g.lo);
        nd.bhi.w = as_float(g.hi - mid);
        nd.alo.w = as_float(mid - g.lo <= leaf ? ~g.lo : mid - 1);
        nd.blo.w = as_float(g.hi - mid <= leaf ? ~mid : mid);
        if (m > leaf && g.plink < 0) T.root = make_int2(id, m);
        Seg b{mid, g.hi, (id << 1) | 1, g.depth + 1, {rlo[0], rlo[1], rlo[2]}, {rhi[0], rhi[1], rhi[2]}};
        Seg a{g.lo, mid, (id << 1) | 0, g.depth + 1, {llo[0], llo[1], llo[2]}, {lhi[0], lhi[1], lhi[2]}};
        stack.push_back(b);
        stack.push_back(a);
    }
    child_boxes(T);
}

// ---------------------------------------------------------------------------
// bounds-checked memory of one lane
// ---------------------------------------------------------------------------
[[noreturn]] void oob(const char* what, long long idx, long long size) {
    throw std::out_of_range(std::string(what) + " index " + std::to_string(idx) + " outside [0, " +
                            std::to_string(size) + ")");
}

template <int S>
struct LaneMemHost {
    const HostTree* T;
    int* ring_r;
    float* ring_dd;
    std::vector<int2>* spill;
    int spill_depth;
    bool turn(bool) const { return true; }
    void fetch4(bool scanning, int first, int node, int last, float4* v) const {
        for (int u = 0; u < 4; ++u) {
            const int o = std::min(u, last);
            if (scanning) {
                const long long i = (long long)first + o;
                if (i < 0 || i >= T->n) oob("point", i, T->n);
                v[u] = T->pts4[i];
            } else {
                if (node < 0 || node >= (int)T->nodes.size()) oob("node", node, (long long)T->nodes.size());
                const Node& nd = T->nodes[node];
                v[u] = o == 0 ? nd.alo : (o == 1 ? nd.ahi : (o == 2 ? nd.blo : nd.bhi));
            }
        }
    }
    int ring_ref(int j) const {
        if (j < 0 || j >= S) oob("ring", j, S);
        return ring_r[j];
    }
    float ring_d(int j) const {
        if (j < 0 || j >= S) oob("ring", j, S);
        return ring_dd[j];
    }
    void ring_put(int j, int r, float d) const {
        if (j < 0 || j >= S) oob("ring", j, S);
        ring_r[j] = r;
        ring_dd[j] = d;
    }
    int2 spill_get(int l) const {
        if (l < 0 || l >= spill_depth) oob("spill", l, spill_depth);
        return (*spill)[l];
    }
    void spill_put(int l, int2 e) const {
        if (l < 0 || l >= spill_depth) oob("spill", l, spill_depth);
        (*spill)[l] = e;
    }
};

struct PtHost {
    const HostTree* T;
    float4 operator()(int64_t i) const {
        if (i < 0 || i >= T->n) oob("re-rank point", i, T->n);
        return T->pts4[i];
    }
};

// One query through the batched form's per-lane path: inline rounds with radius
// R0 * mult^r (the last unbounded), then the guard-band test and the re-rank.
// status: 0 written, 1 handed to the exact kernel (band overflow or stack), 2
// the id guard tripped.
template <int C>
int run_query(const HostTree& T, const float* q, int k, float R0, float mult, int rounds, int stack_depth,
              uint32_t* ids, int64_t* stats) {
    constexpr int G = GUARD, S = 8;
    if (k + G > C) throw std::invalid_argument("k + 3 exceeds the list capacity");
    int ring_r[S] = {0};
    float ring_d[S] = {0};
    std::vector<int2> spill(std::max(stack_depth - S, 0));
    const LaneMemHost<S> mem{&T, ring_r, ring_d, &spill, (int)spill.size()};
    const int max_depth = S + (int)spill.size();
    F32Query<C, G> Q;
    bool overflow = false, resolved = false;
    float R = R0;
    uint64_t visits = 0, tests = 0;
    for (int round = 0;; ++round) {
        const bool last = round >= rounds - 1;
        const float T0 = last ? INFINITY : R * R;
        Q.start(q[0], q[1], q[2], T0, k, true);
        overflow = false;
        if (T.root.y > 0) stats[2] += lane_walk<C, G, S, true>(Q, T.root, mem, T.n + 64, max_depth, overflow, visits,
                                                               tests);
        resolved = last || (Q.kth() < INFINITY && Q.kth() <= T0);
        if (resolved || overflow) break;
        R *= mult;
    }
    stats[0] += (int64_t)visits;
    stats[1] += (int64_t)tests;
    const float kth = Q.kth();
    const float tf = band(kth);
    if (!resolved || overflow || Q.ev_min <= tf) return 1;
    double dd[C];
    uint32_t id[C];
    bool bad = false;
    rerank_sort<C>(Q.L.d, Q.L.s, Q.qx, Q.qy, Q.qz, tf, 0, T.n, PtHost{&T}, dd, id, bad);
    for (int j = 0; j < k; ++j) ids[j] = id[j];
    return bad ? 2 : 0;
}

// ---------------------------------------------------------------------------
// Barnes-Hut: the host build of bh.hip (mass refit + bh_walk), and per body an
// FP64 evaluation of the same accepted set with a bound on its distance to the
// direct sum.
//
// The bound (Plummer law, phi(r) = 1/sqrt(|r|^2 + eps^2), body force
// G m f(r) with f = -grad phi).  For an accepted child with FP32 centre c and
// mass M_f, R = c - x_i, d = |R|, bodies j at c + delta_j:
//   sum_j m_j f(R + delta_j) - M_f f(R)
//     = (M - M_f) f(R) + Df(R) [sum_j m_j delta_j] + sum_j m_j rem_j,
//   |rem_j| <= (1/2) |delta_j|^2 sup_t ||D^3 phi(R + t delta_j)||.
// Along a line, 1/|(r, eps) + t (u, 0)| (a point of R^4) expands in Legendre
// polynomials, so the n-th directional derivative of phi at r is
// n! P_n(cos) / (|r|^2 + eps^2)^{(n+1)/2} with |P_n| <= 1; the norm of a
// symmetric multilinear form is its largest value on equal unit vectors, so
//   ||D^2 phi(r)|| <= 2 / (|r|^2+eps^2)^{3/2},  ||D^3 phi(r)|| <= 6 / (|r|^2+eps^2)^2,
// and |f(R)| <= 1/(d^2+eps^2).  With Q = sum m_j |delta_j|^2, b = max |delta_j|
// (bodies of positive mass) and |R + t delta_j| >= max(0, d - b):
//   error <= |M - M_f| / (d^2+eps^2) + 2 |sum m_j delta_j| / (d^2+eps^2)^{3/2}
//            + 3 Q / (max(0, d - b)^2 + eps^2)^2,
// summed over the body's accepted children (times G).  Bodies summed one by
// one contribute no approximation error.
// ---------------------------------------------------------------------------
struct BHMemHost {
    const HostTree* T;
    const std::vector<BHNode>* bhn;
    const std::vector<float>* mass_;
    Node node(int i) const {
        if (i < 0 || i >= (int)T->nodes.size()) oob("node", i, (long long)T->nodes.size());
        return T->nodes[i];
    }
    BHNode bh(int i) const {
        if (i < 0 || i >= (int)bhn->size()) oob("bh node", i, (long long)bhn->size());
        return (*bhn)[i];
    }
    float4 pt(int i) const {
        if (i < 0 || i >= T->n) oob("body", i, T->n);
        return T->pts4[i];
    }
    float mass(int i) const {
        if (i < 0 || i >= T->n) oob("mass", i, T->n);
        return (*mass_)[i];
    }
};

// the GPU refit's values: each child slot holds mass_merge over the child's
// binary Karras subtree, leaves (sorted position, mass) first
void mass_refit(const HostTree& T, const std::vector<float>& mass, std::vector<BHNode>& bh) {
    const int n = T.n;
    bh.assign(std::max(n - 1, 1), BHNode{});
    if (n < 2) return;
    std::vector<int> order(n - 1);
    for (int i = 0; i < n - 1; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return T.ranges[a].y - T.ranges[a].x < T.ranges[b].y - T.ranges[b].x;
    });
    std::vector<float4> val(n - 1);
    auto leafv = [&](int i) {
        const float4 p = T.pts4[i];
        return make_float4(p.x, p.y, p.z, mass[i]);
    };
    for (int i : order) {
        const int lo = T.ranges[i].x, hi = T.ranges[i].y;
        const int gamma = lo + as_int(T.nodes[i].ahi.w) - 1;
        const float4 a = lo == gamma ? leafv(lo) : val[gamma];
        const float4 b = hi == gamma + 1 ? leafv(gamma + 1) : val[gamma + 1];
        bh[i].a = a;
        bh[i].b = b;
        val[i] = mass_merge(a, b);
    }
}

// FP64 moments of one child's bodies about its FP32 centre
struct ChildMoments {
    double m = 0, s1[3] = {0, 0, 0}, q = 0, b = 0;
};

void child_moments(const HostTree& T, const std::vector<float>& mass, const std::vector<BHNode>& bh,
                   std::vector<ChildMoments>& out) {
    const int n = T.n;
    out.assign(2 * (size_t)std::max(n - 1, 0), ChildMoments{});
    for (int i = 0; i < n - 1; ++i) {
        const int lo = T.ranges[i].x, hi = T.ranges[i].y, cA = as_int(T.nodes[i].ahi.w);
        for (int side = 0; side < 2; ++side) {
            const float4 c = side ? bh[i].b : bh[i].a;
            const int f = side ? lo + cA : lo, l = side ? hi : lo + cA - 1;
            ChildMoments& M = out[2 * (size_t)i + side];
            for (int t = f; t <= l; ++t) {
                const double m = mass[t];
                const double dx = (double)T.pts4[t].x - c.x, dy = (double)T.pts4[t].y - c.y,
                             dz = (double)T.pts4[t].z - c.z;
                const double r2 = dx * dx + dy * dy + dz * dz;
                M.m += m;
                M.s1[0] += m * dx;
                M.s1[1] += m * dy;
                M.s1[2] += m * dz;
                M.q += m * r2;
                if (m > 0) M.b = std::max(M.b, std::sqrt(r2));
            }
        }
    }
}

// one FP64 force term G m d / (|d|^2 + eps^2)^{3/2} (or the reference law) at
// offset d; adds it to acc and its magnitude to absum
void term64(bool ref_law, double G, double e2, double m, double dx, double dy, double dz, double* acc,
            double& absum) {
    const double d2 = (dx * dx + dy * dy) + dz * dz;
    double w;
    if (ref_law) {
        if (!(d2 > 1e-10)) return;
        w = G * m / (d2 * std::sqrt(d2 + e2));
    } else {
        const double r2 = d2 + e2;
        w = G * m / (r2 * std::sqrt(r2));
    }
    acc[0] += w * dx;
    acc[1] += w * dy;
    acc[2] += w * dz;
    absum += std::fabs(w) * std::sqrt(d2);
}

template <bool REFLAW>
void bh_body(const HostTree& T, const std::vector<BHNode>& bh, const std::vector<float>& mass,
             const std::vector<ChildMoments>& mom, int i, float theta2, float eps, float G, float* acc32,
             double* acc64, double& bound, double& absum, BHCounts& cnt, bool& ok) {
    const float4 me = T.pts4[i];
    BHAcc<REFLAW> A;
    A.x = me.x;
    A.y = me.y;
    A.z = me.z;
    A.eps2 = eps * eps;
    A.G = G;
    const double e2 = (double)eps * (double)eps, G64 = G;
    acc64[0] = acc64[1] = acc64[2] = 0.0;
    bound = 0.0;
    absum = 0.0;
    const BHMemHost M{&T, &bh, &mass};
    auto on_far = [&](int node, int side, const float4& c) {
        const double Rx = (double)c.x - me.x, Ry = (double)c.y - me.y, Rz = (double)c.z - me.z;
        term64(REFLAW, G64, e2, c.w, Rx, Ry, Rz, acc64, absum);
        const ChildMoments& cm = mom[2 * (size_t)node + side];
        const double d = std::sqrt(Rx * Rx + Ry * Ry + Rz * Rz);
        const double den = d * d + e2;
        const double s1 = std::sqrt(cm.s1[0] * cm.s1[0] + cm.s1[1] * cm.s1[1] + cm.s1[2] * cm.s1[2]);
        const double g = std::max(0.0, d - cm.b);
        bound += std::fabs(G64) * (std::fabs(cm.m - (double)c.w) / den + 2.0 * s1 / (den * std::sqrt(den)) +
                                   3.0 * cm.q / ((g * g + e2) * (g * g + e2)));
    };
    auto on_leaf = [&](int f, int m) {
        for (int t = f; t < f + m; ++t)
            term64(REFLAW, G64, e2, mass[t], (double)T.pts4[t].x - me.x, (double)T.pts4[t].y - me.y,
                   (double)T.pts4[t].z - me.z, acc64, absum);
    };
    ok = bh_walk<EXACT_STACK, true>(A, T.root, theta2, T.n + 64, M, cnt, on_far, on_leaf);
    acc32[0] = A.ax.value();
    acc32[1] = A.ay.value();
    acc32[2] = A.az.value();
    if (REFLAW) bound = NAN;
}

thread_local std::string g_msg;

}  // namespace

extern "C" {

__attribute__((visibility("default"))) const char* gprt_host_last_error(void) { return g_msg.c_str(); }

// The LBVH of one cloud as lbvh.hip builds it: keys (n, sorted), perm (n),
// nodes (16 floats per internal node), root (2 ints).  Any output may be NULL.
__attribute__((visibility("default"))) int gprt_host_build_tree(const float* pts, int64_t n, int leaf_size,
                                                                uint64_t* keys, uint32_t* perm, float* nodes,
                                                                int32_t* root) {
    if (n < 0 || n > 0x7FFFFFF0 || leaf_size < 1 || leaf_size > 32) return 1;
    HostTree T;
    build_tree(pts, (int)n, leaf_size, T);
    if (keys) std::memcpy(keys, T.keys.data(), 8 * (size_t)n);
    if (perm) std::memcpy(perm, T.perm.data(), 4 * (size_t)n);
    if (nodes && !T.nodes.empty()) std::memcpy(nodes, T.nodes.data(), sizeof(Node) * T.nodes.size());
    if (root) {
        root[0] = T.root.x;
        root[1] = T.root.y;
    }
    return 0;
}

// Either tree with its refit links: mode 0 the LBVH, 1 the SAH tree of
// build_tree_sah.  keys / perm / nodes / root as gprt_host_build_tree; parent
// (2n-1 ints) and ranges (2 ints per internal node; [0, -1] for an unused id).
// Any output may be NULL.
__attribute__((visibility("default"))) int gprt_host_build_tree_ex(const float* pts, int64_t n, int leaf_size,
                                                                   int mode, uint64_t* keys, uint32_t* perm,
                                                                   float* nodes, int32_t* root, int32_t* parent,
                                                                   int32_t* ranges) {
    g_msg.clear();
    try {
        if (n < 0 || n > 0x7FFFFFF0 || leaf_size < 1 || leaf_size > 32 || mode < 0 || mode > 1)
            throw std::invalid_argument("bad argument");
        HostTree T;
        if (mode == 1) build_tree_sah(pts, (int)n, leaf_size, T);
        else build_tree(pts, (int)n, leaf_size, T);
        if (keys) std::memcpy(keys, T.keys.data(), 8 * (size_t)n);
        if (perm) std::memcpy(perm, T.perm.data(), 4 * (size_t)n);
        if (nodes && !T.nodes.empty() && n > 1) std::memcpy(nodes, T.nodes.data(), sizeof(Node) * (size_t)(n - 1));
        if (root) {
            root[0] = T.root.x;
            root[1] = T.root.y;
        }
        if (parent && n > 1) std::memcpy(parent, T.parent.data(), sizeof(int) * (size_t)(2 * n - 1));
        if (ranges && n > 1) std::memcpy(ranges, T.ranges.data(), sizeof(int2) * (size_t)(n - 1));
    } catch (const std::exception& e) {
        g_msg = e.what();
        return 1;
    }
    return 0;
}

// The batched form's per-lane path for each query (see run_query) with a list
// of `cap` slots (8, 16, 24, 32, 40, 48, 64 or 67).  ids_out: nq*k (rows of
// status != 0 are left untouched); status_out: nq; stats_out[3]: node visits,
// point tests, walk trips.  Returns 0, or 1 with gprt_host_last_error() naming
// the first out-of-bounds access or bad argument.
// tree_mode: 0 the LBVH, 1 the SAH tree.
__attribute__((visibility("default"))) int gprt_host_lane_knn_ex(const float* pts, int64_t n, int leaf_size,
                                                                 int tree_mode, const float* qs, int64_t nq, int k,
                                                                 int cap, float radius0, float mult, int rounds,
                                                                 int stack_depth, uint32_t* ids_out,
                                                                 int32_t* status_out, int64_t* stats_out) {
    g_msg.clear();
    try {
        if (n < 0 || n > 0x7FFFFFF0 || nq < 0 || k < 1 || rounds < 1 || stack_depth < 8 || !(mult > 1.0f) ||
            tree_mode < 0 || tree_mode > 1)
            throw std::invalid_argument("bad argument");
        HostTree T;
        if (tree_mode == 1) build_tree_sah(pts, (int)n, leaf_size, T);
        else build_tree(pts, (int)n, leaf_size, T);
        int64_t st[3] = {0, 0, 0};
        for (int64_t i = 0; i < nq; ++i) {
            const float* q = qs + 3 * i;
            uint32_t* o = ids_out + i * k;
            int r;
            switch (cap) {
                case 8: r = run_query<8>(T, q, k, radius0, mult, rounds, stack_depth, o, st); break;
                case 16: r = run_query<16>(T, q, k, radius0, mult, rounds, stack_depth, o, st); break;
                case 24: r = run_query<24>(T, q, k, radius0, mult, rounds, stack_depth, o, st); break;
                case 32: r = run_query<32>(T, q, k, radius0, mult, rounds, stack_depth, o, st); break;
                case 40: r = run_query<40>(T, q, k, radius0, mult, rounds, stack_depth, o, st); break;
                case 48: r = run_query<48>(T, q, k, radius0, mult, rounds, stack_depth, o, st); break;
                case 64: r = run_query<64>(T, q, k, radius0, mult, rounds, stack_depth, o, st); break;
                case 67: r = run_query<67>(T, q, k, radius0, mult, rounds, stack_depth, o, st); break;
                default: throw std::invalid_argument("cap must be 8, 16, 24, 32, 40, 48, 64 or 67");
            }
            status_out[i] = r;
        }
        if (stats_out) std::memcpy(stats_out, st, sizeof(st));
    } catch (const std::exception& e) {
        g_msg = e.what();
        return 1;
    }
    return 0;
}

__attribute__((visibility("default"))) int gprt_host_lane_knn(const float* pts, int64_t n, int leaf_size,
                                                              const float* qs, int64_t nq, int k, int cap,
                                                              float radius0, float mult, int rounds, int stack_depth,
                                                              uint32_t* ids_out, int32_t* status_out,
                                                              int64_t* stats_out) {
    return gprt_host_lane_knn_ex(pts, n, leaf_size, 0, qs, nq, k, cap, radius0, mult, rounds, stack_depth, ids_out,
                                 status_out, stats_out);
}

// Barnes-Hut for the bodies `sel` (original indices; NULL = all n) as
// gprt_hip_bh_forces computes them with this leaf size: acc32 (nsel*3, the
// FP32 results the kernel's operations give), acc64 (nsel*3, the same
// accepted children and bodies summed in FP64, eps^2 taken in FP64), bound
// (nsel: the bound on |acc64 - direct sum| derived above; NaN for the
// reference law), absum (nsel: sum of the magnitudes of the summed terms),
// stats[3] (node visits, bodies summed, children accepted over the selected
// bodies).  Any output but acc32 may be NULL.  Returns 0, or 1 with
// gprt_host_last_error() naming an out-of-bounds access, a walk over its step
// or stack limit, or a bad argument.
__attribute__((visibility("default"))) int gprt_host_bh(const float* pm, int64_t n, int leaf_size, float theta,
                                                        float eps, float G, int ref_law, const int64_t* sel,
                                                        int64_t nsel, float* acc32, double* acc64, double* bound,
                                                        double* absum, int64_t* stats) {
    g_msg.clear();
    try {
        if (n < 1 || n > 0x7FFFFFF0 || leaf_size < 1 || leaf_size > 32 || nsel < 0 || !acc32 || !(theta >= 0.0f))
            throw std::invalid_argument("bad argument");
        if (!(eps > 0.0f) && !ref_law) throw std::invalid_argument("eps must be > 0 for the Plummer law");
        std::vector<float> xyz(3 * (size_t)n);
        for (int64_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) xyz[3 * i + a] = pm[4 * i + a];
        HostTree T;
        build_tree(xyz.data(), (int)n, leaf_size, T);
        std::vector<float> mass(n);
        std::vector<int> inv(n);
        for (int64_t i = 0; i < n; ++i) {
            mass[i] = pm[4 * (size_t)T.perm[i] + 3];
            inv[T.perm[i]] = (int)i;
        }
        std::vector<BHNode> bh;
        mass_refit(T, mass, bh);
        std::vector<ChildMoments> mom;
        child_moments(T, mass, bh, mom);
        const float theta2 = theta * theta;
        BHCounts cnt;
        const int64_t ns = sel ? nsel : n;
        for (int64_t s = 0; s < ns; ++s) {
            const int64_t o = sel ? sel[s] : s;
            if (o < 0 || o >= n) throw std::out_of_range("selected body outside [0, n)");
            double a64[3], bd, ab;
            bool ok;
            if (ref_law)
                bh_body<true>(T, bh, mass, mom, inv[o], theta2, eps, G, acc32 + 3 * s, a64, bd, ab, cnt, ok);
            else
                bh_body<false>(T, bh, mass, mom, inv[o], theta2, eps, G, acc32 + 3 * s, a64, bd, ab, cnt, ok);
            if (!ok) throw std::runtime_error("walk over its step or stack limit");
            if (acc64) std::memcpy(acc64 + 3 * s, a64, sizeof(a64));
            if (bound) bound[s] = bd;
            if (absum) absum[s] = ab;
        }
        if (stats) {
            stats[0] = (int64_t)cnt.visits;
            stats[1] = (int64_t)cnt.bodies;
            stats[2] = (int64_t)cnt.accepted;
        }
    } catch (const std::exception& e) {
        g_msg = e.what();
        return 1;
    }
    return 0;
}

}  // extern "C"
