// gprt_internal.hpp -- types and device helpers shared by the HIP translation
// units of libgprt_hip.so.  Built for gfx950 only (64-lane wavefronts).
//
// Numeric contract (every TU is compiled with -ffp-contract=off, so a product
// and a sum are rounded separately unless an fma is written explicitly):
//   d64(q,p) = ((dx*dx + dy*dy) + dz*dz) in binary64, dx = (double)q.x-(double)p.x
//   d32(q,p) = fma(dz,dz, fma(dy,dy, dx*dx)) in binary32, dx = q.x - p.x
//   box32(q,B) / box64(q,B): the same formula applied to q - clamp(q, lo, hi).
// Rounding is monotone, so box32(q,B) <= d32(q,p) and box64(q,B) <= d64(q,p)
// for every point p inside B: pruning a box whose value exceeds a threshold
// never drops a point whose own value is within it.
#pragma once
#include <array>
#include <mutex>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <vector>

#include "../../include/gprt_hip.h"

namespace gprt {

// ---------------------------------------------------------------------------
// error plumbing: every entry point returns an int status; the message of the
// last failure on the calling thread is kept for gprt_hip_last_error().
// ---------------------------------------------------------------------------
void set_error(int code, const std::string& msg);
int fail(int code, const std::string& msg);

#define GPRT_HIP_TRY(expr)                                                         \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess)                                                      \
            return ::gprt::fail(GPRT_ERR_HIP, std::string(#expr) + ": " +          \
                                                  hipGetErrorString(_e));          \
    } while (0)

#define GPRT_TRY(expr)            \
    do {                          \
        int _rc = (expr);         \
        if (_rc != GPRT_OK) return _rc; \
    } while (0)

// ---------------------------------------------------------------------------
// device buffer: grow-only, owned by a context or a tree
// ---------------------------------------------------------------------------
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool borrowed = false;  // a view into memory the buffer does not own (never freed, never grown)
    int reserve(size_t need);
    void release();
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};
// While set on the calling thread, a DevBuf::reserve that would allocate
// fails with GPRT_ERR_CAPACITY instead (the allocation-free entry points).
void set_no_alloc(bool on);

// ---------------------------------------------------------------------------
// LBVH layout in HBM
// ---------------------------------------------------------------------------
// One internal node = the boxes of its two children plus their references,
// 64 bytes, read with four 16-byte loads.
//   lo[c].w = child reference (int bits): >= 0 internal node index,
//             < 0 leaf bucket whose first sorted point is ~ref
//   hi[c].w = number of points under the child (int bits)
struct __align__(16) Node {
    float4 alo, ahi, blo, bhi;
};

// 4-wide view of the same tree for the k-NN walk: entry i holds the children
// of binary node i's two children (its grandchildren), or the child itself
// when that child is a leaf bucket; a walk from a root that moves from an
// entry to one of its internal slots only ever reads entries two binary levels
// apart, so it sees a 4-ary tree of half the depth.  Structure of arrays over
// the four slots, 128 bytes, read with eight 16-byte loads:
//   ref[j] >= 0: internal binary node (its own 4-wide entry), < 0: leaf bucket
//   at sorted point ~ref[j]; cnt[j] = points under the slot, 0 = empty slot.
struct __align__(16) Node4 {
    float4 lox, loy, loz, hix, hiy, hiz;
    int4 ref, cnt;
};

// Points in Morton order: (x, y, z, id) with id = index of the point in the
// caller's cloud (int bits in .w).
struct Tree {
    int n = 0;          // total points (all clouds)
    int nclouds = 0;
    int leaf_size = 8;  // subtrees with <= leaf_size points are scanned linearly
    DevBuf pts4;        // float4[n]
    DevBuf nodes;       // Node[max(n-1,1)]
    DevBuf nodes4;      // Node4[max(n-1,1)], the 4-wide view used by the k-NN rounds
    DevBuf roots;       // int2[nclouds]: (ref, count) of each cloud's root
    DevBuf cloud_off;   // int64[nclouds+1] point offsets per cloud (device)
    DevBuf cloud_bbox;  // float[6*nclouds]: per-cloud (min xyz, max xyz)
    DevBuf key_bbox;    // float[6*nclouds]: box the sort keys are normalised to
    DevBuf perm;        // uint32[n]: sorted position -> input index
    DevBuf orig4;       // float4[n]: the points in input order (x, y, z, 0), read by the radius samplers
    DevBuf codes;       // uint64[n]: sorted keys (cloud << 3*axis_bits | morton)
    DevBuf parent;      // int[n-1 + n]: parent of internal nodes, then of points
    DevBuf flags;       // int[n-1]: refit arrival counters
    DevBuf ranges;      // int2[n-1]: sorted-point range [lo, hi] of each internal node
    DevBuf ready;       // int[n]: refit work list (+ its counter at the end)
    DevBuf bh;          // Barnes-Hut per-node data (allocated on demand)
    std::vector<int64_t> h_cloud_off;
    float h_bbox[6] = {0, 0, 0, 0, 0, 0};
    bool build_wide = true;  // also build the 4-wide view (k-NN trees)
    bool sah = false;        // split by binned SAH (sah.hip) instead of the Karras hierarchy
    int key_bits = 30;  // significant bits of the sort key
    int axis_bits = 10; // Morton bits per axis
    int max_depth_bound = 64;
    uint64_t gen = 0;   // build number, different for every build in the process
};

// the device buffers of a tree, in a fixed order (a context's pool of reusable
// tree buffers follows it)
constexpr int TREE_BUFS = 15;
inline std::array<DevBuf*, TREE_BUFS> tree_bufs(Tree& T) {
    return {&T.pts4, &T.nodes,  &T.nodes4, &T.roots,  &T.cloud_off, &T.cloud_bbox, &T.key_bbox, &T.perm,
            &T.orig4, &T.codes, &T.parent, &T.flags,  &T.ranges,    &T.ready,      &T.bh};
}

struct Ctx;

// ---------------------------------------------------------------------------
// numeric helpers.  The ones the per-lane k-NN walk uses are __host__
// __device__, so the walk also compiles for the host (a CPU build of the same
// code, with bounds-checked memory, for tests).
// ---------------------------------------------------------------------------
#define GPRT_HD __host__ __device__ __forceinline__

GPRT_HD int as_int(float f) { return __builtin_bit_cast(int, f); }
GPRT_HD float as_float(int i) { return __builtin_bit_cast(float, i); }

// median of three (one v_med3_f32 on the device)
GPRT_HD float med3f(float a, float b, float c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fmed3f(a, b, c);
#else
    return fmaxf(fminf(a, b), fminf(fmaxf(a, b), c));
#endif
}

GPRT_HD float d2f(float qx, float qy, float qz, float px, float py, float pz) {
    float dx = qx - px, dy = qy - py, dz = qz - pz;
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

GPRT_HD double d2d(float qx, float qy, float qz, float px, float py, float pz) {
    double dx = (double)qx - (double)px;
    double dy = (double)qy - (double)py;
    double dz = (double)qz - (double)pz;
    return (dx * dx + dy * dy) + dz * dz;
}

// largest float <= v for a squared distance v >= 0 (one ulp down when the
// conversion rounded up): a box bound kept in FP32 never prunes more than its
// FP64 value would
__device__ __forceinline__ float f64_lower_f32(double v) {
    float f = (float)v;
    if ((double)f > v) f = __int_as_float(__float_as_int(f) - 1);
    return f;
}

// clamp to [lo, hi] for lo <= hi (every box here is built from points): the
// median of the three, one v_med3_f32
GPRT_HD float clampf(float v, float lo, float hi) { return med3f(v, lo, hi); }

GPRT_HD float box2f(float qx, float qy, float qz, const float4& lo, const float4& hi) {
    return d2f(qx, qy, qz, clampf(qx, lo.x, hi.x), clampf(qy, lo.y, hi.y), clampf(qz, lo.z, hi.z));
}

GPRT_HD double box2d(float qx, float qy, float qz, const float4& lo, const float4& hi) {
    return d2d(qx, qy, qz, clampf(qx, lo.x, hi.x), clampf(qy, lo.y, hi.y), clampf(qz, lo.z, hi.z));
}

// Guard band for the FP32 pass.  |d32 - D| <= 6u D and |d64 - D| <= 6*2^-53 D
// for the exact D, so every point whose (d64, id) ranks within the true top k
// has d32 <= T * (1+6u)(1+e64)/((1-6u)(1-e64)) < T * (1 + 16u), T being the
// k-th smallest d32.  BAND = 1 + 2^-19 = 1 + 32u leaves a factor-2 margin; the
// additive term covers flushed subnormal squares (each < FLT_MIN).
#define GPRT_BAND 1.0000019073486328125f
#define GPRT_BAND_ABS 1e-36f
GPRT_HD float band(float t) { return __builtin_fmaf(t, GPRT_BAND, GPRT_BAND_ABS); }

// Squared FP32 distance from q to the farthest corner of a box: per axis
// max(|q - lo|, |q - hi|), then the d2f sum.
GPRT_HD float boxmax2f(float qx, float qy, float qz, const float4& lo, const float4& hi) {
    const float mx = fmaxf(fabsf(qx - lo.x), fabsf(qx - hi.x));
    const float my = fmaxf(fabsf(qy - lo.y), fabsf(qy - hi.y));
    const float mz = fmaxf(fabsf(qz - lo.z), fabsf(qz - hi.z));
    return __builtin_fmaf(mz, mz, __builtin_fmaf(my, my, mx * mx));
}

// Inner band: boxmax2f(q, box) <= inner_band(fl32(r*r)) proves d64(q, p) <=
// (double)r*(double)r for every point p of the box.  Per axis |q - p| <= M,
// the exact farthest-corner offset, and fl32 rounding is monotone, so
// boxmax2f >= sum M^2 * (1-u)^5 (u = 2^-24: the offsets, a product and two
// fma roundings) while d64(p) <= sum M^2 * (1 + 2^-53)^5.  With s = fl32(r*r)
// <= r^2 (1+u), a threshold s * (1 - 2^-19) = s * (1 - 32u) leaves a factor
// above 4 over the 7u needed; the subtracted 1e-36 covers flushed subnormal
// squares.  A threshold <= 0 (tiny radii) proves nothing.
#define GPRT_INNER 0.9999980926513671875f
GPRT_HD float inner_band(float s) { return __builtin_fmaf(s, GPRT_INNER, -GPRT_BAND_ABS); }

// wave helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
    int l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Wave-aggregated append to a global counter: every active lane with `pred`
// gets a unique slot; returns the slot (or -1 when pred is false).  Must be
// called by all lanes of the wave that are live at this point.
__device__ __forceinline__ int wave_append(unsigned* counter, bool pred) {
    uint64_t m = __ballot(pred);
    if (m == 0) return -1;
    int leader = __ffsll((unsigned long long)m) - 1;
    unsigned base = 0;
    if (__lane_id() == leader) base = atomicAdd(counter, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    return pred ? (int)(base + __popcll(m & lanemask_lt())) : -1;
}

// ---------------------------------------------------------------------------
// internal entry points between translation units
// ---------------------------------------------------------------------------
// axis_bits: Morton bits per axis of the sort keys (0: the widest that fits)
int tree_build(Ctx* ctx, const float* d_pts, const int64_t* h_cloud_off, int nclouds, int leaf_size, Tree* t,
               int axis_bits = 0);
int morton_codes_dev(Ctx* ctx, const float* d_pts, int n, uint32_t* d_codes, float* d_bbox6);
// the points in input order as float4 (x, y, z, 0), as the radius samplers read them
int launch_input_points(hipStream_t s, const float* pts, float4* out, int n);
// the SAH hierarchy over the Morton-sorted points of t (sah.hip): reorders
// pts4 / perm / codes inside each cloud, writes node references and counts,
// parent links, ranges and roots; the refit fills the boxes afterwards
int sah_build(Ctx* ctx, Tree* t);
int radix_sort_pairs_u64(Ctx* ctx, uint64_t* keys, uint32_t* vals, int n, int begin_bit, int end_bit);
// with res_keys/res_vals the result may stay in the context's sort temporaries
// (see radix_sort_impl); without them it is copied back into keys/vals
int radix_sort_pairs_u32(Ctx* ctx, uint32_t* keys, uint32_t* vals, int n, int begin_bit, int end_bit,
                         uint32_t** res_keys = nullptr, uint32_t** res_vals = nullptr);
size_t radix_sort_hist_bytes(int n);

// the runtime context
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 256;
    DevBuf ws[8];          // grow-only scratch slots used by the drivers
    DevBuf sort_tmp[4];    // radix sort ping-pong + histograms
    DevBuf counters;       // small device counters
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    // side stream (default priority) for work that overlaps the main stream:
    // fork/join events without timing
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    gprt_hip_run_stats stats{};
    // cached graph of the batched k-NN pipeline
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    uint64_t graph_key = 0;
    bool kernel_timed = false;
    // pinned host copy of the counter buffer plus one float (the sampled
    // initial radius), filled by copies enqueued behind each call's work
    unsigned* hstage = nullptr;
    // the device buffers of the last tree destroyed on this context, taken
    // over by the next build (rebuilding a tree per frame allocates nothing
    // once the sizes settle)
    DevBuf tree_pool[TREE_BUFS];
};

}  // namespace gprt

// the opaque handles of the C ABI; every call that uses a context (directly,
// through one of its trees, or through a legacy pipeline) holds its mutex for
// the whole call, so calls on one context from several threads run one at a
// time; recursive because some entry points call others
struct gprt_hip_tree;
struct gprt_hip_ctx {
    gprt::Ctx c;
    std::recursive_mutex mu;
    // the trees built on this context and not yet destroyed; destroying the
    // context first leaves them without an owner (they then free their own
    // buffers)
    std::vector<gprt_hip_tree*> trees;
};
struct gprt_hip_tree {
    gprt::Tree t;
    gprt_hip_ctx* owner = nullptr;
};
#define GPRT_CTX_LOCK(h) std::lock_guard<std::recursive_mutex> gprt_ctx_lock_((h)->mu)
