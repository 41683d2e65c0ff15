// api.hip -- the gprt_hip_* C ABI: contexts, trees, and the k-NN drivers.
//
// The batched driver is the device-side restatement of the k_nn! round loop
// (gprt-macros/src/lib.rs:408-604): sampler -> query order -> bounded rounds
// with growing radius -> unbounded last round, with no host round trip between
// rounds.  The whole sequence is enqueued on the context stream and, when
// cfg.use_graph is set, captured once into a hipGraph and replayed.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "gprt_internal.hpp"
#include "knn_internal.hpp"


namespace gprt {

static thread_local std::string g_err_msg;

void set_error(int code, const std::string& msg) {
    (void)code;
    g_err_msg = msg;
}
int fail(int code, const std::string& msg) {
    set_error(code, msg);
    return code;
}

// Device allocations of the library's buffers: counted (live buffers and
// bytes, for leak checks), and failed on purpose when a test asks for it:
// GPRT_HIP_FAULT=alloc makes every one fail, GPRT_HIP_FAULT=alloc:N lets N
// succeed first (also settable with gprt_hip_debug_fail_alloc).
static std::atomic<int64_t> g_alloc_calls{0}, g_live_bufs{0}, g_live_bytes{0};
static std::atomic<int64_t> g_alloc_budget{[] {
    const char* f = std::getenv("GPRT_HIP_FAULT");
    if (!f || std::strncmp(f, "alloc", 5) != 0) return (int64_t)-1;
    return f[5] == ':' ? (int64_t)std::max(0LL, std::atoll(f + 6)) : (int64_t)0;
}()};
static thread_local bool t_no_alloc = false;
void set_no_alloc(bool on) { t_no_alloc = on; }

int DevBuf::reserve(size_t need) {
    if (need <= bytes && ptr) return GPRT_OK;
    if (borrowed || t_no_alloc)
        return fail(GPRT_ERR_CAPACITY, "allocation-free call: a workspace of " + std::to_string(need) +
                                           " bytes does not fit its slot of the scratch array");
    release();
    size_t b = std::max<size_t>(need, 256);
    b = (b + 255) & ~(size_t)255;
    int64_t left = g_alloc_budget.load();
    while (left > 0 && !g_alloc_budget.compare_exchange_weak(left, left - 1)) {}
    if (left == 0) return fail(GPRT_ERR_ALLOC, "hipMalloc of " + std::to_string(b) + " bytes failed (injected)");
    hipError_t e = hipMalloc(&ptr, b);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        (void)hipGetLastError();
        return fail(GPRT_ERR_ALLOC, "hipMalloc of " + std::to_string(b) + " bytes failed");
    }
    bytes = b;
    ++g_alloc_calls;
    ++g_live_bufs;
    g_live_bytes += (int64_t)b;
    return GPRT_OK;
}
void DevBuf::release() {
    if (ptr && !borrowed) {
        (void)hipFree(ptr);
        --g_live_bufs;
        g_live_bytes -= (int64_t)bytes;
    }
    ptr = nullptr;
    bytes = 0;
    borrowed = false;
}

// workspace slots of a context
enum {
    WS_QKEYS = 0,  // uint64[nq]
    WS_QPERM,      // uint32[nq]
    WS_QS4,        // float4[nq]
    WS_LISTA,      // int[nq]
    WS_LISTB,      // int[nq]
    WS_FB,         // int2[nq]: fallback list
    WS_MISC,       // radius/kth/offsets
    WS_BBOXTMP,    // tree-build scratch
};

// extra per-context buffers beyond the 8 generic slots
struct CtxExtra {
    DevBuf q_in, ids_out, d2_out, qcloud, kc, qoff, outbase, flag;
    DevBuf spill;  // traversal-stack entries beyond the LDS ring, int2[depth][lanes]
    // range query
    DevBuf r_counts, r_d2, r_big, r_keys, r_perm, r_qs4, r_sums, r_off;
    // Barnes-Hut
    DevBuf bh_xyz, bh_mass, bh_acc;
    Tree bh_tree;
};
static std::mutex g_extra_mu;
static std::vector<std::pair<Ctx*, CtxExtra*>> g_extra;
static CtxExtra* extra(Ctx* c) {
    std::lock_guard<std::mutex> lk(g_extra_mu);
    for (auto& p : g_extra)
        if (p.first == c) return p.second;
    auto* e = new CtxExtra();
    g_extra.push_back({c, e});
    return e;
}
static void drop_extra(Ctx* c) {
    std::lock_guard<std::mutex> lk(g_extra_mu);
    for (size_t i = 0; i < g_extra.size(); ++i)
        if (g_extra[i].first == c) {
            CtxExtra* e = g_extra[i].second;
            for (DevBuf* b : {&e->q_in, &e->ids_out, &e->d2_out, &e->qcloud, &e->kc, &e->qoff, &e->outbase, &e->flag,
                              &e->spill,
                              &e->r_counts, &e->r_d2, &e->r_big, &e->r_keys, &e->r_perm, &e->r_qs4, &e->r_sums,
                              &e->r_off, &e->bh_xyz, &e->bh_mass, &e->bh_acc})
                b->release();
            Tree& T = e->bh_tree;
            for (DevBuf* b : {&T.pts4, &T.nodes, &T.nodes4, &T.roots, &T.cloud_off, &T.cloud_bbox, &T.key_bbox, &T.perm,
                              &T.orig4, &T.codes, &T.parent, &T.flags, &T.ranges, &T.ready, &T.bh})
                b->release();
            delete e;
            g_extra.erase(g_extra.begin() + i);
            return;
        }
}

// ---------------------------------------------------------------------------
// input validation: finite and |x| <= 1e18 (keeps every d^2 finite in FP32)
// ---------------------------------------------------------------------------
__global__ void validate_kernel(const float* __restrict__ v, int64_t m, unsigned* flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    for (; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        float x = v[i];
        bad |= !(fabsf(x) <= 1e18f);
    }
    if (__ballot(bad) && __lane_id() == 0) atomicOr(flag, 1u);
}

static void launch_validate(hipStream_t s, const float* v, int64_t m, unsigned* flag) {
    if (m <= 0) return;
    int blocks = (int)std::min<int64_t>((m + 255) / 256, 1024);
    hipLaunchKernelGGL(validate_kernel, dim3(blocks), dim3(256), 0, s, v, m, flag);
}

// before each k class: its work counters restart and its round 0 holds `size`
__global__ void class_init_kernel(unsigned* counters, unsigned size) {
    const int t = threadIdx.x;
    if (t <= CNT_FALLBACK) counters[t] = t == CNT_ACTIVE ? size : 0u;
}

// clears the call's counters and, for one cloud, writes its query offsets
// zeroes the call's counters (and the query sort's scratch prefix, z2/n2) and
// writes the one-cloud query offsets
__global__ void init_call_kernel(unsigned* counters, int n, int64_t* qoff, int64_t nq, unsigned* z2, int n2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += gridDim.x * blockDim.x) z2[i] = 0u;
    if (blockIdx.x != 0) return;
    for (int i = threadIdx.x; i < n; i += blockDim.x) counters[i] = 0u;
    if (qoff && threadIdx.x == 0) {
        qoff[0] = 0;
        qoff[1] = nq;
    }
}

int range_run(Ctx* ctx, const Tree& T, const float* d_q, int nq, float r, int64_t* d_off, uint32_t* d_ids,
              int64_t capacity, DevBuf& counts_buf, DevBuf& d2_buf, DevBuf& big_buf, DevBuf& keys_buf,
              DevBuf& perm_buf, DevBuf& qs4_buf, DevBuf& sums_buf, int64_t* h_total, const float* d_radii = nullptr, bool count = false, bool time_walk = false);
int bh_run(Ctx* ctx, Tree& T, const float4* d_pm, int n, float theta, float eps, float G, float* d_acc,
           DevBuf& mass_buf, bool count, bool ref_law, bool lane_walk);
int bh_prepare(Ctx* ctx, const float4* d_pm, int n, float* d_xyz, unsigned* flag);

}  // namespace gprt

using namespace gprt;

// ===========================================================================
// library / context
// ===========================================================================
extern "C" int gprt_hip_version(void) { return 10000; }

extern "C" void gprt_hip_debug_alloc_stats(int64_t* calls, int64_t* live_buffers, int64_t* live_bytes) {
    if (calls) *calls = g_alloc_calls.load();
    if (live_buffers) *live_buffers = g_live_bufs.load();
    if (live_bytes) *live_bytes = g_live_bytes.load();
}

extern "C" void gprt_hip_debug_fail_alloc(int64_t after) { g_alloc_budget.store(after < 0 ? -1 : after); }

// a plain streaming copy, 16 bytes per lane and step: reads and writes exactly
// `bytes`, the reference point for scaling hardware byte counters
__global__ __launch_bounds__(256) void stream_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                          size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" int gprt_hip_debug_stream_copy(gprt_hip_ctx* h, void* dst, const void* src, size_t bytes) {
    if (!h || !dst || !src || bytes % 16) return fail(GPRT_ERR_ARG, "null pointer or bytes not a multiple of 16");
    GPRT_CTX_LOCK(h);
    GPRT_HIP_TRY(hipSetDevice(h->c.device));
    const size_t n = bytes / 16;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, (size_t)h->c.num_cus * 8);
    if (n) hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, h->c.stream, (const float4*)src,
                              (float4*)dst, n);
    GPRT_HIP_TRY(hipGetLastError());
    GPRT_HIP_TRY(hipStreamSynchronize(h->c.stream));
    return GPRT_OK;
}

extern "C" const char* gprt_hip_last_error(void) { return g_err_msg.c_str(); }
extern "C" void gprt_hip_clear_error(void) { g_err_msg.clear(); }

extern "C" void gprt_hip_config_default(gprt_hip_config* c) {
    if (!c) return;
    c->sampler = GPRT_SAMPLER_LCG;
    c->radius_heuristic = GPRT_RADIUS_PERCENTILE;
    c->radius_param = 0.95f;
    c->radius_mult = 3.0f;
    c->max_rounds = 2;
    c->traversal = GPRT_TRAV_THREAD;
    c->stack_depth = 64;
    c->guard_slots = 3;
    c->sort_queries = 1;
    c->use_graph = 1;
    c->grid_blocks = 0;
    c->block_threads = 0;
    c->collect_counters = 0;
    c->time_kernels = 0;
    c->lds_stack = 8;
    c->compact_rounds = 0;
}

extern "C" int gprt_hip_num_devices(int* out) {
    if (!out) return fail(GPRT_ERR_ARG, "null out");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *out = n;
    return GPRT_OK;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free k-NN entry point (gprt_hip_knn_noalloc): a
// file-scope device array, one copy per GPU, held until the process ends.  Its
// address on each device is looked up once, when the first context on that
// device is created, and only that address is handed to kernels and copies.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

// with `device` current
static int scratch_setup(int device) {
    if (device < 0 || device >= 64) return fail(GPRT_ERR_ARG, "device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return GPRT_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return fail(GPRT_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch_ptr[device] = ptr;
    return GPRT_OK;
}

extern "C" int gprt_hip_ctx_create(int device, gprt_hip_ctx** out) {
    if (!out) return fail(GPRT_ERR_ARG, "null out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return fail(GPRT_ERR_HIP, "no HIP device available");
    }
    if (device < 0 || device >= n) return fail(GPRT_ERR_ARG, "device index out of range");
    GPRT_HIP_TRY(hipSetDevice(device));
    auto* h = new gprt_hip_ctx();
    Ctx& c = h->c;
    c.device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c.num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c.ev0) != hipSuccess || hipEventCreate(&c.ev1) != hipSuccess ||
        hipEventCreate(&c.ev2) != hipSuccess || hipEventCreate(&c.ev3) != hipSuccess ||
        hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming) != hipSuccess) {
        gprt_hip_ctx_destroy(h);
        return fail(GPRT_ERR_HIP, "stream/event creation failed");
    }
    if (c.counters.reserve(sizeof(unsigned) * CNT_ALL) != GPRT_OK) {
        gprt_hip_ctx_destroy(h);
        return GPRT_ERR_ALLOC;
    }
    if (hipHostMalloc((void**)&c.hstage, sizeof(unsigned) * (CNT_ALL + 1), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c.hstage = nullptr;
        gprt_hip_ctx_destroy(h);
        return fail(GPRT_ERR_ALLOC, "pinned host allocation failed");
    }
    if (int rc = scratch_setup(device); rc != GPRT_OK) {
        gprt_hip_ctx_destroy(h);
        return rc;
    }
    (void)extra(&c);  // the per-context extras exist before any allocation-free call
    *out = h;
    return GPRT_OK;
}

extern "C" int gprt_hip_ctx_destroy(gprt_hip_ctx* h) {
    if (!h) return GPRT_OK;
    {
        // trees that outlive their context keep working and free their own
        // buffers when destroyed
        GPRT_CTX_LOCK(h);
        for (gprt_hip_tree* t : h->trees) t->owner = nullptr;
        h->trees.clear();
    }
    Ctx& c = h->c;
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    if (c.graph_exec) (void)hipGraphExecDestroy(c.graph_exec);
    if (c.graph) (void)hipGraphDestroy(c.graph);
    for (auto& b : c.ws) b.release();
    for (auto& b : c.sort_tmp) b.release();
    c.counters.release();
    for (auto& b : c.tree_pool) b.release();
    if (c.hstage) (void)hipHostFree(c.hstage);
    drop_extra(&c);
    if (c.side) (void)hipStreamSynchronize(c.side);
    for (hipEvent_t e : {c.ev0, c.ev1, c.ev2, c.ev3, c.ev_fork, c.ev_join})
        if (e) (void)hipEventDestroy(e);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    if (c.side) (void)hipStreamDestroy(c.side);
    delete h;
    return GPRT_OK;
}

extern "C" int gprt_hip_ctx_sync(gprt_hip_ctx* h) {
    if (!h) return fail(GPRT_ERR_ARG, "null ctx");
    GPRT_CTX_LOCK(h);
    GPRT_HIP_TRY(hipStreamSynchronize(h->c.stream));
    return GPRT_OK;
}

extern "C" int gprt_hip_device_alloc(gprt_hip_ctx* h, size_t bytes, void** out) {
    if (!h || !out) return fail(GPRT_ERR_ARG, "null ctx/out");
    GPRT_CTX_LOCK(h);
    *out = nullptr;
    GPRT_HIP_TRY(hipSetDevice(h->c.device));
    if (hipMalloc(out, std::max<size_t>(bytes, 1)) != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(GPRT_ERR_ALLOC, "hipMalloc failed");
    }
    return GPRT_OK;
}

extern "C" int gprt_hip_device_free(gprt_hip_ctx* h, void* p) {
    if (!h) return fail(GPRT_ERR_ARG, "null ctx");
    GPRT_CTX_LOCK(h);
    if (!p) return GPRT_OK;
    GPRT_HIP_TRY(hipSetDevice(h->c.device));
    GPRT_HIP_TRY(hipStreamSynchronize(h->c.stream));
    GPRT_HIP_TRY(hipFree(p));
    return GPRT_OK;
}

extern "C" int gprt_hip_copy(gprt_hip_ctx* h, void* dst, const void* src, size_t bytes, int kind) {
    if (!h || (bytes && (!dst || !src))) return fail(GPRT_ERR_ARG, "null argument");
    if (kind < 0 || kind > 2) return fail(GPRT_ERR_ARG, "bad copy kind");
    GPRT_CTX_LOCK(h);
    if (!bytes) return GPRT_OK;
    GPRT_HIP_TRY(hipSetDevice(h->c.device));
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    GPRT_HIP_TRY(hipMemcpyAsync(dst, src, bytes, k, h->c.stream));
    GPRT_HIP_TRY(hipStreamSynchronize(h->c.stream));
    return GPRT_OK;
}

// ===========================================================================
// trees
// ===========================================================================
static int build_common(gprt_hip_ctx* h, const float* pts, const int64_t* off, int nclouds, int leaf_size,
                        int flags, gprt_hip_tree** out) {
    if (!h || !out) return fail(GPRT_ERR_ARG, "null ctx/out");
    *out = nullptr;
    GPRT_CTX_LOCK(h);
    if (nclouds <= 0) return fail(GPRT_ERR_ARG, "nclouds must be positive");
    if (leaf_size < 1 || leaf_size > 32) return fail(GPRT_ERR_ARG, "leaf_size must be in [1, 32]");
    for (int c = 0; c < nclouds; ++c)
        if (off[c + 1] < off[c] || off[0] != 0) return fail(GPRT_ERR_ARG, "cloud offsets must be non-decreasing from 0");
    int64_t n = off[nclouds];
    if (n > 0 && !pts) return fail(GPRT_ERR_ARG, "null points");
    Ctx& c = h->c;
    GPRT_HIP_TRY(hipSetDevice(c.device));
    CtxExtra* x = extra(&c);
    const float* d_pts = pts;
    if (!(flags & GPRT_PTR_DEVICE) && n > 0) {
        GPRT_TRY(x->q_in.reserve(sizeof(float) * 3 * n));
        GPRT_HIP_TRY(hipMemcpyAsync(x->q_in.ptr, pts, sizeof(float) * 3 * n, hipMemcpyHostToDevice, c.stream));
        d_pts = x->q_in.as<float>();
    }
    GPRT_TRY(x->flag.reserve(16));
    GPRT_HIP_TRY(hipMemsetAsync(x->flag.ptr, 0, 16, c.stream));
    launch_validate(c.stream, d_pts, 3 * n, x->flag.as<unsigned>());
    unsigned bad = 0;
    GPRT_HIP_TRY(hipMemcpyAsync(&bad, x->flag.ptr, 4, hipMemcpyDeviceToHost, c.stream));
    GPRT_HIP_TRY(hipStreamSynchronize(c.stream));
    if (bad) return fail(GPRT_ERR_INPUT, "points contain NaN/Inf or |coordinate| > 1e18");
    auto* t = new gprt_hip_tree();
    t->owner = h;
    h->trees.push_back(t);
    static std::atomic<uint64_t> builds{0};
    t->t.gen = ++builds;
    {
        // the buffers of the context's last destroyed tree (grown below if small)
        auto bufs = tree_bufs(t->t);
        for (int i = 0; i < TREE_BUFS; ++i) std::swap(*bufs[i], c.tree_pool[i]);
    }
    GPRT_HIP_TRY(hipEventRecord(c.ev2, c.stream));
    t->t.sah = (flags & GPRT_BUILD_SAH) != 0;
    int rc = tree_build(&c, d_pts, off, nclouds, leaf_size, &t->t, (flags & GPRT_KEYS_REF30) ? 10 : 0);
    if (rc != GPRT_OK) {
        gprt_hip_tree_destroy(t);
        return rc;
    }
    GPRT_HIP_TRY(hipEventRecord(c.ev3, c.stream));
    GPRT_HIP_TRY(hipEventSynchronize(c.ev3));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c.ev2, c.ev3);
    c.stats.ms_build = ms;
    *out = t;
    return GPRT_OK;
}

extern "C" int gprt_hip_tree_build(gprt_hip_ctx* h, const float* pts, int64_t n, int leaf_size, int flags,
                                   gprt_hip_tree** out) {
    if (n < 0) return fail(GPRT_ERR_ARG, "n < 0");
    int64_t off[2] = {0, n};
    return build_common(h, pts, off, 1, leaf_size, flags, out);
}

extern "C" int gprt_hip_tree_build_batched(gprt_hip_ctx* h, const float* pts, const int64_t* cloud_offsets,
                                           int nclouds, int leaf_size, int flags, gprt_hip_tree** out) {
    if (!cloud_offsets) return fail(GPRT_ERR_ARG, "null cloud_offsets");
    return build_common(h, pts, cloud_offsets, nclouds, leaf_size, flags, out);
}

extern "C" int gprt_hip_tree_destroy(gprt_hip_tree* t) {
    if (!t) return GPRT_OK;
    std::unique_lock<std::recursive_mutex> lk;
    if (t->owner) {
        lk = std::unique_lock<std::recursive_mutex>(t->owner->mu);
        (void)hipSetDevice(t->owner->c.device);
    }
    auto bufs = tree_bufs(t->t);
    if (t->owner) {
        // kept for the owner's next build.  Work that still reads the tree may
        // sit on any context's stream, so wait for the device first, as
        // freeing the buffers would
        (void)hipDeviceSynchronize();
        DevBuf* pool = t->owner->c.tree_pool;
        for (int i = 0; i < TREE_BUFS; ++i) {
            pool[i].release();
            std::swap(*bufs[i], pool[i]);
        }
        auto& live = t->owner->trees;
        live.erase(std::remove(live.begin(), live.end(), t), live.end());
    } else {
        for (DevBuf* b : bufs) b->release();
    }
    delete t;
    return GPRT_OK;
}

extern "C" int gprt_hip_tree_info(const gprt_hip_tree* t, gprt_hip_tree_desc* out) {
    if (!t || !out) return fail(GPRT_ERR_ARG, "null tree/out");
    out->n = t->t.n;
    out->nclouds = t->t.nclouds;
    out->leaf_size = t->t.leaf_size;
    for (int a = 0; a < 6; ++a) out->bbox[a] = t->t.h_bbox[a];
    out->key_bits = t->t.key_bits;
    out->axis_bits = t->t.axis_bits;
    return GPRT_OK;
}

extern "C" int gprt_hip_tree_export(gprt_hip_tree* t, uint64_t* keys, uint32_t* perm, float* nodes,
                                    int32_t* roots) {
    if (!t) return fail(GPRT_ERR_ARG, "null tree");
    if (!t->owner) return fail(GPRT_ERR_STATE, "the tree's context was destroyed");
    GPRT_CTX_LOCK(t->owner);
    Ctx& c = t->owner->c;
    Tree& T = t->t;
    hipStream_t s = c.stream;
    GPRT_HIP_TRY(hipSetDevice(c.device));
    if (keys && T.n) GPRT_HIP_TRY(hipMemcpyAsync(keys, T.codes.ptr, 8ull * T.n, hipMemcpyDeviceToHost, s));
    if (perm && T.n) GPRT_HIP_TRY(hipMemcpyAsync(perm, T.perm.ptr, 4ull * T.n, hipMemcpyDeviceToHost, s));
    if (nodes && T.n > 1)
        GPRT_HIP_TRY(hipMemcpyAsync(nodes, T.nodes.ptr, sizeof(Node) * (T.n - 1), hipMemcpyDeviceToHost, s));
    if (roots) GPRT_HIP_TRY(hipMemcpyAsync(roots, T.roots.ptr, 8ull * T.nclouds, hipMemcpyDeviceToHost, s));
    GPRT_HIP_TRY(hipStreamSynchronize(s));
    return GPRT_OK;
}

// ===========================================================================
// k-NN drivers
// ===========================================================================
namespace gprt {

struct KnnPlan {
    const Tree* tree;
    const float* d_q;     // xyz, original order, device
    int64_t nq;
    int nclouds;          // query sets (== tree clouds)
    int kmax;
    int k;                // uniform k (batched: 0)
    const int* d_kc;      // per cloud k or nullptr
    const int64_t* d_qoff;
    const int64_t* d_outbase;
    int64_t max_q_per_cloud;
    gprt_hip_config cfg;
    uint32_t* d_ids;
    double* d_d2;
    // k classes (sorted order groups queries by class)
    int class_begin[N_KCLASS];
    int class_size[N_KCLASS];
    int class_kmax[N_KCLASS];
};

// fills the k-class layout of a plan from host-side per-cloud k and query counts
static void plan_classes(KnnPlan& p, const int* h_k, const int64_t* h_qoff, int nclouds, int k_uniform) {
    for (int c = 0; c < N_KCLASS; ++c) p.class_size[c] = p.class_kmax[c] = 0;
    for (int c = 0; c < nclouds; ++c) {
        const int k = h_k ? h_k[c] : k_uniform;
        const int cl = k_class(k);
        p.class_size[cl] += (int)(h_qoff[c + 1] - h_qoff[c]);
        p.class_kmax[cl] = std::max(p.class_kmax[cl], k);
    }
    int b = 0;
    for (int c = 0; c < N_KCLASS; ++c) {
        p.class_begin[c] = b;
        b += p.class_size[c];
    }
}

static int knn_reserve(Ctx& c, const KnnPlan& p) {
    size_t nq = (size_t)std::max<int64_t>(p.nq, 1);
    GPRT_TRY(c.ws[WS_QKEYS].reserve(8 * nq));
    GPRT_TRY(c.ws[WS_QPERM].reserve(std::max(4 * nq, query_prep_scratch_bytes(p.nclouds, p.cfg.sort_queries))));
    GPRT_TRY(c.ws[WS_QS4].reserve(16 * nq));
    GPRT_TRY(c.ws[WS_LISTA].reserve(4 * nq));
    GPRT_TRY(c.ws[WS_LISTB].reserve(4 * nq));
    GPRT_TRY(c.ws[WS_FB].reserve(8 * nq));
    GPRT_TRY(c.ws[WS_MISC].reserve(sizeof(float) * 101 * std::max(p.nclouds, 1) + 1024));
    CtxExtra* x = extra(&c);
    if (p.nclouds > 1) GPRT_TRY(x->qcloud.reserve(4 * nq));
    // HBM part of the per-lane traversal stacks of the persistent rounds
    if (p.cfg.traversal != GPRT_TRAV_WAVE) {
        size_t bytes = 0;
        for (int cl = 0; cl < N_LANE_CLASSES; ++cl)
            if (p.class_size[cl] > 0) {
                const size_t lanes = (size_t)256 * lane_round_grid(cl, p.cfg.lds_stack, p.cfg.grid_blocks,
                                                                   p.cfg.traversal);
                bytes = std::max(bytes, 8 * lanes * lane_spill_depth(p.cfg.stack_depth, cl, p.cfg.lds_stack));
            }
        if (bytes) GPRT_TRY(x->spill.reserve(bytes));
    }
    // radix sort temporaries for nq keys
    GPRT_TRY(c.sort_tmp[0].reserve(8 * nq));
    GPRT_TRY(c.sort_tmp[1].reserve(4 * nq));
    GPRT_TRY(c.sort_tmp[2].reserve(radix_sort_hist_bytes((int)nq)));
    return GPRT_OK;
}

static int check_config(const gprt_hip_config& cf) {
    if (cf.max_rounds < 0) return fail(GPRT_ERR_ARG, "max_rounds < 0");
    if (cf.traversal < GPRT_TRAV_THREAD || cf.traversal > GPRT_TRAV_WIDE) return fail(GPRT_ERR_ARG, "bad traversal");
    if (cf.stack_depth < 1 || cf.stack_depth > 4096) return fail(GPRT_ERR_ARG, "stack_depth must be in [1, 4096]");
    if (cf.lds_stack < 1 || cf.lds_stack > 32) return fail(GPRT_ERR_ARG, "lds_stack must be in [1, 32]");
    if (cf.guard_slots != 3) return fail(GPRT_ERR_ARG, "guard_slots must be 3 (the lists carry exactly 3)");
    if (cf.grid_blocks < 0 || cf.grid_blocks > (1 << 20)) return fail(GPRT_ERR_ARG, "bad grid_blocks");
    if (cf.sampler != GPRT_SAMPLER_LCG && cf.sampler != GPRT_SAMPLER_STRIDED) return fail(GPRT_ERR_ARG, "bad sampler");
    if (cf.radius_heuristic < GPRT_RADIUS_MAX || cf.radius_heuristic > GPRT_RADIUS_INFINITE)
        return fail(GPRT_ERR_ARG, "bad radius_heuristic");
    if (!(cf.radius_mult > 1.0f) || !(cf.radius_mult < 1e30f)) return fail(GPRT_ERR_ARG, "radius_mult must be > 1");
    return GPRT_OK;
}


static int knn_enqueue(Ctx& c, const KnnPlan& p) {
    hipStream_t s = c.stream;
    const Tree& T = *p.tree;
    const gprt_hip_config& cfg = p.cfg;
    unsigned* counters = c.counters.as<unsigned>();
    unsigned long long* stats64 = (unsigned long long*)(counters + CNT_TOTAL);
    float* kth = c.ws[WS_MISC].as<float>();
    float* radius = kth + 100 * p.nclouds;
    int nq = (int)p.nq;
    CtxExtra* x = extra(&c);
    int nclasses = 0;
    for (int cl = 0; cl < N_KCLASS; ++cl) nclasses += p.class_size[cl] > 0;
    unsigned* sort_prefix = nullptr;
    int sort_prefix_words = 0;
    GPRT_TRY(query_prep_clear_range(&c, p.nclouds, cfg.sort_queries, nq, nclasses > 1, &sort_prefix,
                                    &sort_prefix_words));
    hipLaunchKernelGGL(init_call_kernel, dim3(1 + sort_prefix_words / 4096), dim3(256), 0, s, counters, CNT_ALL,
                       p.nclouds == 1 ? (int64_t*)p.d_qoff : nullptr, p.nq, sort_prefix, sort_prefix_words);
    uint64_t* keys = c.ws[WS_QKEYS].as<uint64_t>();
    uint32_t* perm = c.ws[WS_QPERM].as<uint32_t>();
    float4* qs4 = c.ws[WS_QS4].as<float4>();
    int* qcloud = p.nclouds > 1 ? x->qcloud.as<int>() : nullptr;
    // the radius sampler runs on the side stream, overlapping the query sort;
    // the rounds wait for both.  The longer branch is enqueued first (a
    // captured graph launches its branches in that order): the sort for one
    // cloud, the sampler (100 x 5000 draws per cloud) for several.
    GPRT_HIP_TRY(hipEventRecord(c.ev_fork, s));
    auto sort = [&]() {
        return launch_query_prep(&c, p.d_q, p.d_qoff, p.nclouds, T.key_bbox.as<float>(), cfg.sort_queries, nq, keys,
                                 perm, qs4, counters, p.max_q_per_cloud, qcloud, perm, counters + CNT_TOTAL - 1,
                                 p.d_kc, p.k, nclasses > 1, sort_prefix_words > 0);
    };
    if (p.nclouds == 1) GPRT_TRY(sort());
    GPRT_HIP_TRY(hipStreamWaitEvent(c.side, c.ev_fork, 0));
    GPRT_TRY(launch_sampler(c.side, cfg.sampler, T.orig4.as<float4>(), T.cloud_off.as<int64_t>(), p.d_q, p.d_qoff,
                            p.nclouds, p.kmax, cfg.radius_heuristic, cfg.radius_param, kth, radius));
    if (p.nclouds > 1) GPRT_TRY(sort());
    GPRT_HIP_TRY(hipEventRecord(c.ev_join, c.side));
    GPRT_HIP_TRY(hipStreamWaitEvent(s, c.ev_join, 0));
    int grid = cfg.grid_blocks > 0 ? cfg.grid_blocks : 0;
    int R = cfg.radius_heuristic == GPRT_RADIUS_INFINITE ? 1 : std::min(cfg.max_rounds, MAX_ROUNDS - 1) + 1;
    int* lists[2] = {c.ws[WS_LISTA].as<int>(), c.ws[WS_LISTB].as<int>()};
    RoundArgs ra{};
    ra.nodes = T.nodes.as<Node>();
    ra.nodes4 = T.nodes4.as<Node4>();
    ra.pts = T.pts4.as<float4>();
    ra.roots = T.roots.as<int2>();
    ra.qs = qs4;
    ra.qcloud = qcloud;
    ra.kc = p.d_kc;
    ra.k = p.k;
    ra.radius0 = radius;
    ra.mult = cfg.radius_mult;
    ra.fb_list = c.ws[WS_FB].as<int2>();
    ra.counters = counters;
    ra.stats64 = stats64;
    ra.ids_out = p.d_ids;
    ra.d2_out = p.d_d2;
    ra.out_base = p.d_outbase;
    ra.qoff = p.d_qoff;
    ra.max_steps = T.n + 64;
    ra.spill = x->spill.as<int2>();
    ra.poff = T.cloud_off.as<int64_t>();
    ExactArgs ea{};
    ea.nq = p.nq;
    ea.qs = qs4;
    ea.fb_list = c.ws[WS_FB].as<int2>();
    ea.qcloud = qcloud;
    ea.kc = p.d_kc;
    ea.k = p.k;
    ea.nodes = T.nodes.as<Node>();
    ea.pts = T.pts4.as<float4>();
    ea.roots = T.roots.as<int2>();
    ea.counters = counters;
    ea.stats64 = stats64;
    ea.ids_out = p.d_ids;
    ea.d2_out = p.d_d2;
    ea.out_base = p.d_outbase;
    ea.qoff = p.d_qoff;
    ea.max_steps = T.n + 64;
    const bool time_k = cfg.time_kernels && !cfg.use_graph;
    bool timed = false;
    for (int cl = 0; cl < N_KCLASS; ++cl) {
        const int size = p.class_size[cl];
        if (size == 0) continue;
        if (nclasses > 1)
            hipLaunchKernelGGL(class_init_kernel, dim3(1), dim3(64), 0, s, counters, (unsigned)size);
        ra.base = p.class_begin[cl];
        if (cl == QWAVE_CLASS) {
            // large k: one query per wavefront, exact in binary64, rounds retried
            // inline; nothing falls back
            ra.round = 0;
            ra.rounds = R;
            ra.last = R == 1;
            ra.inline_rounds = 1;
            GPRT_TRY(launch_qwave(s, ra, p.class_kmax[cl], grid, cfg.collect_counters != 0));
            continue;
        }
        ra.spill_depth = lane_spill_depth(cfg.stack_depth, cl, cfg.lds_stack);
        // later rounds inside the first launch (a lane retries its unresolved
        // query with the next radius) unless the schedule asks for one launch
        // per round over the compacted unresolved list; packets always compact
        const int inl = (!cfg.compact_rounds && cfg.traversal != GPRT_TRAV_WAVE) ? 1 : 0;
        ra.rounds = R;
        ra.inline_rounds = inl;
        for (int r = 0; r < (inl ? 1 : R); ++r) {
            ra.round = r;
            ra.last = (r == R - 1);
            ra.in_list = r == 0 ? nullptr : lists[r & 1];
            ra.out_list = lists[(r + 1) & 1];
            const bool tk = r == 0 && time_k && !timed;
            if (tk) GPRT_HIP_TRY(hipEventRecord(c.ev2, s));
            GPRT_TRY(launch_round(s, ra, cl, grid, cfg.traversal, cfg.lds_stack, cfg.collect_counters != 0));
            if (tk) {
                GPRT_HIP_TRY(hipEventRecord(c.ev3, s));
                timed = true;
            }
        }
        ea.mode = 1;
        ea.base = 0;
        ea.nq = size;
        GPRT_TRY(launch_exact(s, ea, p.class_kmax[cl], size, cfg.collect_counters != 0));
    }
    c.kernel_timed = timed;
    return GPRT_OK;
}

static uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

// timed: bracket the call with the ev0/ev1 timing events (read back by
// finish_stats); calls that only enqueue (GPRT_NO_SYNC) leave them out, as an
// event record between two calls keeps the GPU idle while it is processed
static int knn_run(Ctx& c, const KnnPlan& p, bool timed) {
    GPRT_TRY(knn_reserve(c, p));
    hipStream_t s = c.stream;
    if (p.cfg.use_graph) {
        // everything the captured kernels point at: a rebuilt tree or a grown
        // workspace changes a pointer and forces a new capture
        const Tree& T = *p.tree;
        const void* ptrs[] = {T.nodes.ptr, T.nodes4.ptr, T.pts4.ptr, T.roots.ptr, T.orig4.ptr, T.key_bbox.ptr, T.cloud_off.ptr,
                              c.ws[0].ptr, c.ws[1].ptr, c.ws[2].ptr, c.ws[3].ptr, c.ws[4].ptr, c.ws[5].ptr,
                              c.ws[6].ptr, c.sort_tmp[0].ptr, c.sort_tmp[1].ptr, c.sort_tmp[2].ptr,
                              c.counters.ptr, extra(&c)->qcloud.ptr, extra(&c)->spill.ptr};
        uint64_t key = 1469598103934665603ull;
        key = fnv(key, &p, sizeof(KnnPlan));
        key = fnv(key, ptrs, sizeof(ptrs));
        key = fnv(key, &T.n, sizeof(int));
        key = fnv(key, &T.gen, sizeof(T.gen));
        if (!c.graph_exec || c.graph_key != key) {
            if (c.graph_exec) (void)hipGraphExecDestroy(c.graph_exec);
            if (c.graph) (void)hipGraphDestroy(c.graph);
            c.graph_exec = nullptr;
            c.graph = nullptr;
            GPRT_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rc = knn_enqueue(c, p);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s, &g);
            if (rc != GPRT_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(GPRT_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            c.graph = g;
            GPRT_HIP_TRY(hipGraphInstantiate(&c.graph_exec, g, nullptr, nullptr, 0));
            c.graph_key = key;
        }
        if (timed) GPRT_HIP_TRY(hipEventRecord(c.ev0, s));
        GPRT_HIP_TRY(hipGraphLaunch(c.graph_exec, s));
        if (timed) GPRT_HIP_TRY(hipEventRecord(c.ev1, s));
    } else {
        if (timed) GPRT_HIP_TRY(hipEventRecord(c.ev0, s));
        GPRT_TRY(knn_enqueue(c, p));
        if (timed) GPRT_HIP_TRY(hipEventRecord(c.ev1, s));
    }
    return GPRT_OK;
}

// d_radius: the sampled radius of cloud 0 on the device (batched k-NN), else null
static int finish_stats(Ctx& c, int64_t nq, bool read_counters, const float* d_radius = nullptr) {
    // the counters and the radius come back by copies queued behind the work,
    // so the call waits for the GPU once
    if (read_counters)
        GPRT_HIP_TRY(hipMemcpyAsync(c.hstage, c.counters.ptr, sizeof(unsigned) * CNT_ALL, hipMemcpyDeviceToHost,
                                    c.stream));
    if (d_radius)
        GPRT_HIP_TRY(hipMemcpyAsync(c.hstage + CNT_ALL, d_radius, sizeof(float), hipMemcpyDeviceToHost, c.stream));
    GPRT_HIP_TRY(hipStreamSynchronize(c.stream));
    if (d_radius) memcpy(&c.stats.initial_radius, c.hstage + CNT_ALL, sizeof(float));
    float ms = 0;
    if (hipEventElapsedTime(&ms, c.ev0, c.ev1) != hipSuccess) {
        (void)hipGetLastError();
        ms = 0;
    }
    c.stats.ms_query = ms;
    c.stats.ms_total = ms;
    c.stats.queries = nq;
    if (read_counters) {
        const unsigned* h = c.hstage;
        unsigned long long v64[CNT_STATS64];
        memcpy(v64, h + CNT_TOTAL, sizeof(v64));
        c.stats.rerank_reads = (int64_t)v64[2];
        c.stats.wave_steps = (int64_t)v64[4];
        c.stats.lane_steps = (int64_t)v64[5];
        if (c.kernel_timed) {
            float mk = 0;
            if (hipEventElapsedTime(&mk, c.ev2, c.ev3) == hipSuccess) c.stats.ms_kernel = mk;
            c.kernel_timed = false;
        }
        c.stats.rounds_run = 0;
        for (int r = 0; r < 16; ++r) {
            c.stats.active_per_round[r] = h[CNT_ACTIVE + r];
            if (r < MAX_ROUNDS && h[CNT_ACTIVE + r] > 0) c.stats.rounds_run = r + 1;
        }
        c.stats.fallback_queries = h[CNT_FALLBACK];
        c.stats.partial_queries = h[CNT_PARTIAL];
        c.stats.node_visits = (int64_t)v64[0];
        c.stats.point_tests = (int64_t)v64[1];
        if (h[CNT_TOTAL - 1]) return fail(GPRT_ERR_INPUT, "queries contain NaN/Inf or |coordinate| > 1e18");
        if (h[CNT_ERROR]) return fail(GPRT_ERR_STATE, "traversal step guard tripped: corrupt tree");
        if (h[CNT_SORTERR]) return fail(GPRT_ERR_STATE, "query sort: radix sort look-back did not complete");
        if (h[CNT_OFFSETS])
            return fail(GPRT_ERR_ARG, "range: offsets_out does not match these queries and radius (" +
                                          std::to_string(h[CNT_OFFSETS]) + " queries)");
        if (h[CNT_BADID])
            return fail(GPRT_ERR_STATE, "k-NN id guard tripped: " + std::to_string(h[CNT_BADID]) +
                                            " queries re-ranked a candidate outside their cloud");
    }
    return GPRT_OK;
}

}  // namespace gprt


extern "C" int gprt_hip_knn(gprt_hip_ctx* h, gprt_hip_tree* t, const float* queries, int64_t nq, int k,
                            const gprt_hip_config* cfg, uint32_t* ids_out, double* d2_out, int flags) {
    if (!h || !t) return fail(GPRT_ERR_ARG, "null ctx/tree");
    if (k < 1 || k > GPRT_HIP_K_MAX) return fail(GPRT_ERR_ARG, "k must be in [1, " + std::to_string(GPRT_HIP_K_MAX) + "]");
    if (nq < 0 || nq > 0x7FFFFFF0) return fail(GPRT_ERR_ARG, "bad nq");
    if (nq > 0 && (!queries || !ids_out)) return fail(GPRT_ERR_ARG, "null queries/ids_out");
    if (t->t.nclouds != 1) return fail(GPRT_ERR_ARG, "tree has several clouds: use gprt_hip_knn_batched");
    GPRT_CTX_LOCK(h);
    Ctx& c = h->c;
    GPRT_HIP_TRY(hipSetDevice(c.device));
    gprt_hip_config cf;
    if (cfg) cf = *cfg; else gprt_hip_config_default(&cf);
    GPRT_TRY(check_config(cf));
    c.stats = gprt_hip_run_stats{};
    if (nq == 0) return GPRT_OK;
    CtxExtra* x = extra(&c);
    bool dev = flags & GPRT_PTR_DEVICE;
    const float* d_q = queries;
    uint32_t* d_ids = ids_out;
    double* d_d2 = d2_out;
    hipStream_t s = c.stream;
    if (!dev) {
        GPRT_TRY(x->q_in.reserve(12ull * nq));
        GPRT_TRY(x->ids_out.reserve(4ull * nq * k));
        if (d2_out) GPRT_TRY(x->d2_out.reserve(8ull * nq * k));
        GPRT_HIP_TRY(hipMemcpyAsync(x->q_in.ptr, queries, 12ull * nq, hipMemcpyHostToDevice, s));
        d_q = x->q_in.as<float>();
        d_ids = x->ids_out.as<uint32_t>();
        d_d2 = d2_out ? x->d2_out.as<double>() : nullptr;
    }
    GPRT_TRY(x->qoff.reserve(16));
    KnnPlan p{};
    p.tree = &t->t;
    p.d_q = d_q;
    p.nq = nq;
    p.nclouds = 1;
    p.kmax = k;
    p.k = k;
    p.d_kc = nullptr;
    p.d_qoff = x->qoff.as<int64_t>();
    p.d_outbase = nullptr;
    p.max_q_per_cloud = nq;
    p.cfg = cf;
    p.d_ids = d_ids;
    p.d_d2 = d_d2;
    {
        int64_t qo[2] = {0, nq};
        plan_classes(p, nullptr, qo, 1, k);
    }
    GPRT_TRY(knn_run(c, p, !(flags & GPRT_NO_SYNC)));
    if (!dev) {
        GPRT_HIP_TRY(hipMemcpyAsync(ids_out, d_ids, 4ull * nq * k, hipMemcpyDeviceToHost, s));
        if (d2_out) GPRT_HIP_TRY(hipMemcpyAsync(d2_out, d_d2, 8ull * nq * k, hipMemcpyDeviceToHost, s));
    }
    if (flags & GPRT_NO_SYNC) return GPRT_OK;
    return finish_stats(c, nq, true, c.ws[WS_MISC].as<float>() + 100 * (p.nclouds));
}

// ---------------------------------------------------------------------------
// gprt_hip_knn without device allocation: every per-call workspace of the
// batched pipeline (query keys, order and copies, work lists, fallback list,
// sort temporaries, radius slots) is a slot of the scratch array, lent to the
// context's buffers for the duration of the call; the traversal stack is the
// LDS ring only (stack_depth 8, no HBM spill; deeper walks finish in the exact
// kernel) and no graph is captured.  A reservation that would still allocate
// fails the call instead (GPRT_ERR_CAPACITY).
// ---------------------------------------------------------------------------
namespace {
struct ScratchPlan {
    size_t off[12];
    size_t size[12];
    size_t total;
};
// slots: 7 workspaces (WS_QKEYS .. WS_MISC), 3 sort temporaries, qoff
ScratchPlan scratch_plan(int64_t nq) {
    const size_t q = (size_t)std::max<int64_t>(nq, 1);
    const size_t sz[11] = {8 * q, 4 * q, 16 * q, 4 * q, 4 * q, 8 * q, sizeof(float) * 101 + 1024,
                           8 * q, 4 * q, radix_sort_hist_bytes((int)q), 16};
    ScratchPlan p{};
    size_t o = 0;
    for (int i = 0; i < 11; ++i) {
        p.off[i] = o;
        p.size[i] = (sz[i] + 255) & ~(size_t)255;
        o += p.size[i];
    }
    p.total = o;
    return p;
}
}  // namespace

extern "C" int64_t gprt_hip_knn_noalloc_max_queries(void) {
    int64_t lo = 0, hi = (int64_t)NOALLOC_SCRATCH_BYTES;
    while (lo < hi) {
        const int64_t m = (lo + hi + 1) / 2;
        if (scratch_plan(m).total <= NOALLOC_SCRATCH_BYTES) lo = m; else hi = m - 1;
    }
    return lo;
}

extern "C" int gprt_hip_knn_noalloc(gprt_hip_ctx* h, gprt_hip_tree* t, const float* queries, int64_t nq, int k,
                                    const gprt_hip_config* cfg, uint32_t* ids_out, double* d2_out) {
    if (!h || !t) return fail(GPRT_ERR_ARG, "null ctx/tree");
    if (k < 1 || k > GPRT_HIP_K_MAX) return fail(GPRT_ERR_ARG, "k must be in [1, " + std::to_string(GPRT_HIP_K_MAX) + "]");
    if (nq < 0) return fail(GPRT_ERR_ARG, "bad nq");
    if (nq > gprt_hip_knn_noalloc_max_queries())
        return fail(GPRT_ERR_CAPACITY, "gprt_hip_knn_noalloc takes at most " +
                                           std::to_string(gprt_hip_knn_noalloc_max_queries()) + " queries (" +
                                           std::to_string(NOALLOC_SCRATCH_BYTES >> 20) + " MiB scratch array)");
    if (nq > 0 && (!queries || !ids_out)) return fail(GPRT_ERR_ARG, "null queries/ids_out");
    if (t->t.nclouds != 1) return fail(GPRT_ERR_ARG, "tree has several clouds");
    GPRT_CTX_LOCK(h);
    Ctx& c = h->c;
    GPRT_HIP_TRY(hipSetDevice(c.device));
    gprt_hip_config cf;
    if (cfg) cf = *cfg; else gprt_hip_config_default(&cf);
    cf.use_graph = 0;
    cf.stack_depth = 8;  // no HBM spill: the LDS ring is the whole stack
    GPRT_TRY(check_config(cf));
    c.stats = gprt_hip_run_stats{};
    if (nq == 0) return GPRT_OK;
    unsigned char* base = (unsigned char*)g_scratch_ptr[c.device];
    if (!base) return fail(GPRT_ERR_STATE, "scratch array not set up on this device");
    CtxExtra* x = extra(&c);
    DevBuf* slots[11] = {&c.ws[WS_QKEYS], &c.ws[WS_QPERM], &c.ws[WS_QS4], &c.ws[WS_LISTA], &c.ws[WS_LISTB],
                         &c.ws[WS_FB], &c.ws[WS_MISC], &c.sort_tmp[0], &c.sort_tmp[1], &c.sort_tmp[2], &x->qoff};
    DevBuf saved[11];
    const ScratchPlan sp = scratch_plan(nq);
    for (int i = 0; i < 11; ++i) {
        saved[i] = *slots[i];
        slots[i]->ptr = base + sp.off[i];
        slots[i]->bytes = sp.size[i];
        slots[i]->borrowed = true;
    }
    set_no_alloc(true);
    KnnPlan p{};
    p.tree = &t->t;
    p.d_q = queries;
    p.nq = nq;
    p.nclouds = 1;
    p.kmax = k;
    p.k = k;
    p.d_qoff = x->qoff.as<int64_t>();
    p.max_q_per_cloud = nq;
    p.cfg = cf;
    p.d_ids = ids_out;
    p.d_d2 = d2_out;
    {
        int64_t qo[2] = {0, nq};
        plan_classes(p, nullptr, qo, 1, k);
    }
    int rc = knn_run(c, p, true);
    if (rc == GPRT_OK) rc = finish_stats(c, nq, true, c.ws[WS_MISC].as<float>() + 100);
    set_no_alloc(false);
    for (int i = 0; i < 11; ++i) *slots[i] = saved[i];
    return rc;
}

extern "C" int gprt_hip_knn_plain(gprt_hip_ctx* h, gprt_hip_tree* t, const float* queries, int64_t nq, int k,
                                  uint32_t* ids_out, double* d2_out, int flags) {
    if (!h || !t) return fail(GPRT_ERR_ARG, "null ctx/tree");
    GPRT_CTX_LOCK(h);
    if (k < 1 || k > GPRT_HIP_K_MAX) return fail(GPRT_ERR_ARG, "k must be in [1, " + std::to_string(GPRT_HIP_K_MAX) + "]");
    if (nq < 0 || nq > 0x7FFFFFF0) return fail(GPRT_ERR_ARG, "bad nq");
    if (nq > 0 && (!queries || !ids_out)) return fail(GPRT_ERR_ARG, "null queries/ids_out");
    if (t->t.nclouds != 1) return fail(GPRT_ERR_ARG, "plain form takes a single-cloud tree");
    Ctx& c = h->c;
    GPRT_HIP_TRY(hipSetDevice(c.device));
    c.stats = gprt_hip_run_stats{};
    if (nq == 0) return GPRT_OK;
    CtxExtra* x = extra(&c);
    bool dev = flags & GPRT_PTR_DEVICE;
    hipStream_t s = c.stream;
    const float* d_q = queries;
    uint32_t* d_ids = ids_out;
    double* d_d2 = d2_out;
    if (!dev) {
        GPRT_TRY(x->q_in.reserve(12ull * nq));
        GPRT_TRY(x->ids_out.reserve(4ull * nq * k));
        if (d2_out) GPRT_TRY(x->d2_out.reserve(8ull * nq * k));
        GPRT_HIP_TRY(hipMemcpyAsync(x->q_in.ptr, queries, 12ull * nq, hipMemcpyHostToDevice, s));
        d_q = x->q_in.as<float>();
        d_ids = x->ids_out.as<uint32_t>();
        d_d2 = d2_out ? x->d2_out.as<double>() : nullptr;
    }
    const Tree& T = t->t;
    unsigned* counters = c.counters.as<unsigned>();
    ExactArgs ea{};
    ea.mode = 0;
    ea.nq = nq;
    ea.q_xyz = d_q;
    ea.k = k;
    ea.nodes = T.nodes.as<Node>();
    ea.pts = T.pts4.as<float4>();
    ea.roots = T.roots.as<int2>();
    ea.counters = counters;
    ea.stats64 = (unsigned long long*)(counters + CNT_TOTAL);
    ea.ids_out = d_ids;
    ea.d2_out = d_d2;
    ea.max_steps = T.n + 64;
    bool count = (flags & GPRT_COLLECT_COUNTERS) != 0;
    GPRT_TRY(launch_fill_u32(s, counters, CNT_TOTAL + 2 * CNT_STATS64, 0u));
    launch_validate(s, d_q, 3 * nq, counters + CNT_TOTAL - 1);
    const bool timed = !(flags & GPRT_NO_SYNC);
    if (timed) GPRT_HIP_TRY(hipEventRecord(c.ev0, s));
    if (k <= 64) {
        GPRT_TRY(launch_exact(s, ea, k, (int)nq, count));
    } else {
        // beyond the per-lane list: one query per wavefront, same FP64 order
        RoundArgs ra{};
        ra.nodes = ea.nodes;
        ra.pts = ea.pts;
        ra.roots = ea.roots;
        ra.k = k;
        ra.last = 1;
        ra.rounds = 1;
        ra.counters = counters;
        ra.stats64 = ea.stats64;
        ra.ids_out = d_ids;
        ra.d2_out = d_d2;
        ra.max_steps = ea.max_steps;
        ra.poff = T.cloud_off.as<int64_t>();
        ra.q_xyz = d_q;
        ra.nq_plain = nq;
        GPRT_TRY(launch_qwave(s, ra, k, 0, count));
    }
    if (timed) GPRT_HIP_TRY(hipEventRecord(c.ev1, s));
    if (!dev) {
        GPRT_HIP_TRY(hipMemcpyAsync(ids_out, d_ids, 4ull * nq * k, hipMemcpyDeviceToHost, s));
        if (d2_out) GPRT_HIP_TRY(hipMemcpyAsync(d2_out, d_d2, 8ull * nq * k, hipMemcpyDeviceToHost, s));
    }
    if (flags & GPRT_NO_SYNC) return GPRT_OK;
    return finish_stats(c, nq, true);
}

extern "C" int gprt_hip_knn_batched(gprt_hip_ctx* h, gprt_hip_tree* t, const float* queries,
                                    const int64_t* query_offsets, const int* k_per_cloud,
                                    const gprt_hip_config* cfg, uint32_t* ids_out, double* d2_out, int flags) {
    if (!h || !t || !query_offsets || !k_per_cloud) return fail(GPRT_ERR_ARG, "null argument");
    GPRT_CTX_LOCK(h);
    const Tree& T = t->t;
    int nc = T.nclouds;
    if (query_offsets[0] != 0) return fail(GPRT_ERR_ARG, "query_offsets[0] must be 0");
    int kmax = 1;
    std::vector<int64_t> outbase(nc + 1, 0);
    int64_t maxq = 0;
    for (int c = 0; c < nc; ++c) {
        if (query_offsets[c + 1] < query_offsets[c]) return fail(GPRT_ERR_ARG, "query_offsets must be non-decreasing");
        if (k_per_cloud[c] < 1 || k_per_cloud[c] > GPRT_HIP_K_MAX)
            return fail(GPRT_ERR_ARG, "k must be in [1, " + std::to_string(GPRT_HIP_K_MAX) + "]");
        kmax = std::max(kmax, k_per_cloud[c]);
        outbase[c + 1] = outbase[c] + (query_offsets[c + 1] - query_offsets[c]) * k_per_cloud[c];
        maxq = std::max(maxq, query_offsets[c + 1] - query_offsets[c]);
    }
    int64_t nq = query_offsets[nc], nout = outbase[nc];
    if (nq > 0x7FFFFFF0) return fail(GPRT_ERR_ARG, "too many queries");
    if (nq > 0 && (!queries || !ids_out)) return fail(GPRT_ERR_ARG, "null queries/ids_out");
    Ctx& c = h->c;
    GPRT_HIP_TRY(hipSetDevice(c.device));
    gprt_hip_config cf;
    if (cfg) cf = *cfg; else gprt_hip_config_default(&cf);
    GPRT_TRY(check_config(cf));
    c.stats = gprt_hip_run_stats{};
    if (nq == 0) return GPRT_OK;
    CtxExtra* x = extra(&c);
    hipStream_t s = c.stream;
    bool dev = flags & GPRT_PTR_DEVICE;
    const float* d_q = queries;
    uint32_t* d_ids = ids_out;
    double* d_d2 = d2_out;
    if (!dev) {
        GPRT_TRY(x->q_in.reserve(12ull * nq));
        GPRT_TRY(x->ids_out.reserve(4ull * nout));
        if (d2_out) GPRT_TRY(x->d2_out.reserve(8ull * nout));
        GPRT_HIP_TRY(hipMemcpyAsync(x->q_in.ptr, queries, 12ull * nq, hipMemcpyHostToDevice, s));
        d_q = x->q_in.as<float>();
        d_ids = x->ids_out.as<uint32_t>();
        d_d2 = d2_out ? x->d2_out.as<double>() : nullptr;
    }
    GPRT_TRY(x->qoff.reserve(8ull * (nc + 1)));
    GPRT_TRY(x->outbase.reserve(8ull * (nc + 1)));
    GPRT_TRY(x->kc.reserve(4ull * nc));
    GPRT_HIP_TRY(hipMemcpyAsync(x->qoff.ptr, query_offsets, 8ull * (nc + 1), hipMemcpyHostToDevice, s));
    GPRT_HIP_TRY(hipMemcpyAsync(x->outbase.ptr, outbase.data(), 8ull * (nc + 1), hipMemcpyHostToDevice, s));
    GPRT_HIP_TRY(hipMemcpyAsync(x->kc.ptr, k_per_cloud, 4ull * nc, hipMemcpyHostToDevice, s));
    KnnPlan p{};
    p.tree = &T;
    p.d_q = d_q;
    p.nq = nq;
    p.nclouds = nc;
    p.kmax = kmax;
    p.k = 0;
    p.d_kc = x->kc.as<int>();
    p.d_qoff = x->qoff.as<int64_t>();
    p.d_outbase = x->outbase.as<int64_t>();
    p.max_q_per_cloud = maxq;
    p.cfg = cf;
    p.d_ids = d_ids;
    p.d_d2 = d_d2;
    plan_classes(p, k_per_cloud, query_offsets, nc, 0);
    GPRT_TRY(knn_run(c, p, !(flags & GPRT_NO_SYNC)));
    if (!dev) {
        GPRT_HIP_TRY(hipMemcpyAsync(ids_out, d_ids, 4ull * nout, hipMemcpyDeviceToHost, s));
        if (d2_out) GPRT_HIP_TRY(hipMemcpyAsync(d2_out, d_d2, 8ull * nout, hipMemcpyDeviceToHost, s));
    }
    if (flags & GPRT_NO_SYNC) return GPRT_OK;
    return finish_stats(c, nq, true, c.ws[WS_MISC].as<float>() + 100 * (p.nclouds));
}

extern "C" int gprt_hip_stats(gprt_hip_ctx* h, gprt_hip_run_stats* out) {
    if (!h || !out) return fail(GPRT_ERR_ARG, "null ctx/out");
    GPRT_CTX_LOCK(h);
    *out = h->c.stats;
    return GPRT_OK;
}

// ===========================================================================
// pieces exposed for parity tests
// ===========================================================================
extern "C" int gprt_hip_morton_codes(gprt_hip_ctx* h, const float* pts, int64_t n, uint32_t* codes_out,
                                     float* bbox_out, int flags) {
    if (!h || (n > 0 && (!pts || !codes_out))) return fail(GPRT_ERR_ARG, "null argument");
    GPRT_CTX_LOCK(h);
    if (n < 0 || n > 0x7FFFFFF0) return fail(GPRT_ERR_ARG, "bad n");
    Ctx& c = h->c;
    GPRT_HIP_TRY(hipSetDevice(c.device));
    CtxExtra* x = extra(&c);
    hipStream_t s = c.stream;
    bool dev = flags & GPRT_PTR_DEVICE;
    const float* d_p = pts;
    uint32_t* d_codes = codes_out;
    if (!dev) {
        GPRT_TRY(x->q_in.reserve(12ull * std::max<int64_t>(n, 1)));
        GPRT_TRY(x->ids_out.reserve(4ull * std::max<int64_t>(n, 1)));
        if (n) GPRT_HIP_TRY(hipMemcpyAsync(x->q_in.ptr, pts, 12ull * n, hipMemcpyHostToDevice, s));
        d_p = x->q_in.as<float>();
        d_codes = x->ids_out.as<uint32_t>();
    }
    GPRT_TRY(c.ws[WS_MISC].reserve(64));
    float* d_bbox = c.ws[WS_MISC].as<float>();
    GPRT_TRY(morton_codes_dev(&c, d_p, (int)n, d_codes, d_bbox));
    if (!dev && n) GPRT_HIP_TRY(hipMemcpyAsync(codes_out, d_codes, 4ull * n, hipMemcpyDeviceToHost, s));
    if (bbox_out) GPRT_HIP_TRY(hipMemcpyAsync(bbox_out, d_bbox, 24, dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    GPRT_HIP_TRY(hipStreamSynchronize(s));
    return GPRT_OK;
}

extern "C" int gprt_hip_sample_radius(gprt_hip_ctx* h, const float* pts, int64_t n, const float* queries,
                                      int64_t nq, int sampler, int heuristic, float param, int k,
                                      float* radius_out, int flags) {
    if (!h || !radius_out) return fail(GPRT_ERR_ARG, "null argument");
    GPRT_CTX_LOCK(h);
    if (n < 0 || nq < 0 || (n > 0 && !pts) || (nq > 0 && !queries)) return fail(GPRT_ERR_ARG, "bad arrays");
    if (sampler != GPRT_SAMPLER_LCG && sampler != GPRT_SAMPLER_STRIDED) return fail(GPRT_ERR_ARG, "bad sampler");
    if (heuristic < 0 || heuristic > 3) return fail(GPRT_ERR_ARG, "bad heuristic");
    if (sampler == GPRT_SAMPLER_STRIDED && k < 1) return fail(GPRT_ERR_ARG, "k must be positive");
    if (n > (int64_t)0x7FFFFFF0) return fail(GPRT_ERR_ARG, "too many points");
    // the samplers draw input indices and read the points in input order as
    // float4 (one 16-byte load per draw); no tree is needed
    Ctx& c = h->c;
    GPRT_HIP_TRY(hipSetDevice(c.device));
    hipStream_t s = c.stream;
    CtxExtra* x = extra(&c);
    const float* d_pts = pts;
    if (!(flags & GPRT_PTR_DEVICE) && n > 0) {
        GPRT_TRY(x->q_in.reserve(sizeof(float) * 3 * n));
        GPRT_HIP_TRY(hipMemcpyAsync(x->q_in.ptr, pts, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s));
        d_pts = x->q_in.as<float>();
    }
    GPRT_TRY(x->flag.reserve(16));
    GPRT_HIP_TRY(hipMemsetAsync(x->flag.ptr, 0, 16, s));
    launch_validate(s, d_pts, 3 * n, x->flag.as<unsigned>());
    unsigned bad = 0;
    GPRT_HIP_TRY(hipMemcpyAsync(&bad, x->flag.ptr, 4, hipMemcpyDeviceToHost, s));
    GPRT_HIP_TRY(hipStreamSynchronize(s));
    if (bad) return fail(GPRT_ERR_INPUT, "points contain NaN/Inf or |coordinate| > 1e18");
    // orig4 (16 B per point) then the one cloud's offsets {0, n}
    GPRT_TRY(c.sort_tmp[0].reserve(16ull * (size_t)std::max<int64_t>(n, 1) + 64));
    float4* orig4 = c.sort_tmp[0].as<float4>();
    int64_t* poff = (int64_t*)(c.sort_tmp[0].as<char>() + 16ull * (size_t)std::max<int64_t>(n, 1));
    const int64_t po[2] = {0, n};
    GPRT_HIP_TRY(hipMemcpyAsync(poff, po, 16, hipMemcpyHostToDevice, s));
    GPRT_TRY(launch_input_points(s, d_pts, orig4, (int)n));
    int rc = GPRT_OK;    const float* d_q = queries;
    DevBuf qtmp;
    if (!(flags & GPRT_PTR_DEVICE) && nq > 0) {
        rc = qtmp.reserve(12ull * nq);
        if (rc == GPRT_OK && hipMemcpyAsync(qtmp.ptr, queries, 12ull * nq, hipMemcpyHostToDevice, s) != hipSuccess)
            rc = fail(GPRT_ERR_HIP, "copy failed");
        d_q = qtmp.as<float>();
    }
    if (rc == GPRT_OK) rc = x->qoff.reserve(16);
    if (rc == GPRT_OK) rc = c.ws[WS_MISC].reserve(sizeof(float) * 101 + 64);
    if (rc == GPRT_OK) {
        int64_t qo[2] = {0, nq};
        float* kth = c.ws[WS_MISC].as<float>();
        float* rad = kth + 100;
        (void)hipMemcpyAsync(x->qoff.ptr, qo, 16, hipMemcpyHostToDevice, s);
        if (n > 0 && nq > 0)
            rc = launch_sampler(s, sampler, orig4, poff, d_q, x->qoff.as<int64_t>(), 1, k, heuristic, param, kth,
                                rad);
        else
            rc = launch_sampler(s, sampler, nullptr, nullptr, nullptr, nullptr, 1, k, GPRT_RADIUS_INFINITE,
                                param, kth, rad);
        float r = 0;
        if (rc == GPRT_OK && hipMemcpyAsync(&r, rad, 4, hipMemcpyDeviceToHost, s) == hipSuccess &&
            hipStreamSynchronize(s) == hipSuccess) {
            if (!(n > 0 && nq > 0)) {
                // no samples: the reference keeps its default radius
                r = sampler == GPRT_SAMPLER_LCG ? 10.0f
                                                : (heuristic == GPRT_RADIUS_FIXED ? param : 1.0f);
            }
            *radius_out = r;
        } else if (rc == GPRT_OK) {
            rc = fail(GPRT_ERR_HIP, "sampler readback failed");
        }
    }
    (void)hipStreamSynchronize(s);
    qtmp.release();
    return rc;
}

extern "C" int gprt_hip_sort_pairs(gprt_hip_ctx* h, uint64_t* keys, uint32_t* vals, int64_t n, int begin_bit,
                                   int end_bit, int flags) {
    if (!h || (n > 0 && (!keys || !vals))) return fail(GPRT_ERR_ARG, "null argument");
    GPRT_CTX_LOCK(h);
    if (n < 0 || n >= (1 << 30) || begin_bit < 0 || end_bit > 64 || begin_bit > end_bit)
        return fail(GPRT_ERR_ARG, "bad size/bit range (n < 2^30)");
    Ctx& c = h->c;
    GPRT_HIP_TRY(hipSetDevice(c.device));
    hipStream_t s = c.stream;
    bool dev = flags & GPRT_PTR_DEVICE;
    DevBuf kb, vb;
    uint64_t* dk = keys;
    uint32_t* dv = vals;
    if (!dev && n > 0) {
        GPRT_TRY(kb.reserve(8ull * n));
        GPRT_TRY(vb.reserve(4ull * n));
        GPRT_HIP_TRY(hipMemcpyAsync(kb.ptr, keys, 8ull * n, hipMemcpyHostToDevice, s));
        GPRT_HIP_TRY(hipMemcpyAsync(vb.ptr, vals, 4ull * n, hipMemcpyHostToDevice, s));
        dk = kb.as<uint64_t>();
        dv = vb.as<uint32_t>();
    }
    unsigned* sort_err = c.counters.as<unsigned>() + CNT_SORTERR;
    GPRT_TRY(launch_fill_u32(s, sort_err, 1, 0u));
    int rc = radix_sort_pairs_u64(&c, dk, dv, (int)n, begin_bit, end_bit);
    if (rc == GPRT_OK && !dev && n > 0) {
        if (hipMemcpyAsync(keys, dk, 8ull * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(vals, dv, 4ull * n, hipMemcpyDeviceToHost, s) != hipSuccess)
            rc = fail(GPRT_ERR_HIP, "copy back failed");
    }
    unsigned serr = 0;
    if (rc == GPRT_OK && hipMemcpyAsync(&serr, sort_err, 4, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = fail(GPRT_ERR_HIP, "copy back failed");
    if (hipStreamSynchronize(s) != hipSuccess && rc == GPRT_OK) rc = fail(GPRT_ERR_HIP, "sync failed");
    if (rc == GPRT_OK && serr) rc = fail(GPRT_ERR_STATE, "radix sort look-back did not complete");
    kb.release();
    vb.release();
    return rc;
}

// ===========================================================================
// fixed-radius range query
// ===========================================================================
extern "C" int gprt_hip_range(gprt_hip_ctx* h, gprt_hip_tree* t, const float* queries, int64_t nq, float radius,
                              int64_t* offsets_out, uint32_t* ids_out, int64_t ids_capacity, int flags) {
    if (!h || !t || !offsets_out) return fail(GPRT_ERR_ARG, "null ctx/tree/offsets_out");
    GPRT_CTX_LOCK(h);
    if (nq < 0 || nq > 0x7FFFFFF0) return fail(GPRT_ERR_ARG, "bad nq");
    if (nq > 0 && !queries) return fail(GPRT_ERR_ARG, "null queries");
    if (!(radius >= 0.0f && radius <= 1e18f)) return fail(GPRT_ERR_ARG, "radius must be finite and >= 0");
    if (t->t.nclouds != 1) return fail(GPRT_ERR_ARG, "range query takes a single-cloud tree");
    if (ids_out && ids_capacity < 0) return fail(GPRT_ERR_ARG, "negative capacity");
    Ctx& c = h->c;
    GPRT_HIP_TRY(hipSetDevice(c.device));
    CtxExtra* x = extra(&c);
    hipStream_t s = c.stream;
    const bool dev = flags & GPRT_PTR_DEVICE;
    c.stats = gprt_hip_run_stats{};
    if (nq == 0) {
        int64_t z = 0;
        if (dev) GPRT_HIP_TRY(hipMemcpy(offsets_out, &z, 8, hipMemcpyHostToDevice));
        else offsets_out[0] = 0;
        return GPRT_OK;
    }
    const float* d_q = queries;
    if (!dev) {
        GPRT_TRY(x->q_in.reserve(12ull * nq));
        GPRT_HIP_TRY(hipMemcpyAsync(x->q_in.ptr, queries, 12ull * nq, hipMemcpyHostToDevice, s));
        d_q = x->q_in.as<float>();
    }
    int64_t* d_off = offsets_out;
    if (!dev) {
        GPRT_TRY(x->r_off.reserve(8ull * (nq + 1)));
        d_off = x->r_off.as<int64_t>();
    }
    GPRT_TRY(x->flag.reserve(16));
    GPRT_HIP_TRY(hipMemsetAsync(x->flag.ptr, 0, 16, s));
    launch_validate(s, d_q, 3 * nq, x->flag.as<unsigned>());
    unsigned bad = 0;
    GPRT_HIP_TRY(hipMemcpyAsync(&bad, x->flag.ptr, 4, hipMemcpyDeviceToHost, s));
    GPRT_HIP_TRY(hipStreamSynchronize(s));
    if (bad) return fail(GPRT_ERR_INPUT, "queries contain NaN/Inf or |coordinate| > 1e18");
    const Tree& T = t->t;
    int64_t total = 0;
    if (!ids_out) {
        GPRT_HIP_TRY(hipEventRecord(c.ev0, s));
        GPRT_TRY(range_run(&c, T, d_q, (int)nq, radius, d_off, nullptr, 0, x->r_counts, x->r_d2, x->r_big, x->r_keys,
                           x->r_perm, x->r_qs4, x->r_sums, &total, nullptr, (flags & GPRT_COLLECT_COUNTERS) != 0,
                           (flags & GPRT_TIME_KERNELS) != 0));
        GPRT_HIP_TRY(hipEventRecord(c.ev1, s));
        if (!dev) GPRT_HIP_TRY(hipMemcpyAsync(offsets_out, d_off, 8ull * (nq + 1), hipMemcpyDeviceToHost, s));
    } else {
        if (!dev) {
            total = offsets_out[nq];
            GPRT_HIP_TRY(hipMemcpyAsync(d_off, offsets_out, 8ull * (nq + 1), hipMemcpyHostToDevice, s));
        } else {
            GPRT_HIP_TRY(hipMemcpyAsync(&total, d_off + nq, 8, hipMemcpyDeviceToHost, s));
            GPRT_HIP_TRY(hipStreamSynchronize(s));
        }
        if (total > ids_capacity) return fail(GPRT_ERR_CAPACITY, "ids_capacity is smaller than offsets[nq]");
        uint32_t* d_ids = ids_out;
        if (!dev) {
            GPRT_TRY(x->ids_out.reserve(4ull * std::max<int64_t>(total, 1)));
            d_ids = x->ids_out.as<uint32_t>();
        }
        GPRT_HIP_TRY(hipEventRecord(c.ev0, s));
        GPRT_TRY(range_run(&c, T, d_q, (int)nq, radius, d_off, d_ids, total, x->r_counts, x->r_d2, x->r_big,
                           x->r_keys, x->r_perm, x->r_qs4, x->r_sums, nullptr, nullptr,
                           (flags & GPRT_COLLECT_COUNTERS) != 0, (flags & GPRT_TIME_KERNELS) != 0));
        GPRT_HIP_TRY(hipEventRecord(c.ev1, s));
        if (!dev && total) GPRT_HIP_TRY(hipMemcpyAsync(ids_out, d_ids, 4ull * total, hipMemcpyDeviceToHost, s));
    }
    return finish_stats(c, nq, true);
}

// ===========================================================================
// Barnes-Hut
// ===========================================================================
extern "C" int gprt_hip_bh_forces(gprt_hip_ctx* h, const float* pos_mass, int64_t n, float theta, float eps, float G,
                                  int leaf_size, float* acc_out, int flags) {
    if (!h) return fail(GPRT_ERR_ARG, "null ctx");
    GPRT_CTX_LOCK(h);
    if (n < 0 || n > 0x7FFFFFF0) return fail(GPRT_ERR_ARG, "bad n");
    if (n > 0 && (!pos_mass || !acc_out)) return fail(GPRT_ERR_ARG, "null pos_mass/acc_out");
    if (!(theta >= 0.0f && theta <= 1e6f) || !(eps >= 0.0f && eps <= 1e18f) || !(fabsf(G) <= 1e30f))
        return fail(GPRT_ERR_ARG, "theta, eps and G must be finite (theta, eps >= 0)");
    // Plummer softening with eps = 0 divides 0 by 0 on every body's own term
    if (eps == 0.0f && !(flags & GPRT_BH_REFERENCE_LAW))
        return fail(GPRT_ERR_ARG, "eps must be > 0 for the Plummer law (GPRT_BH_REFERENCE_LAW has no softening floor)");
    if (leaf_size < 1 || leaf_size > 32) return fail(GPRT_ERR_ARG, "leaf_size must be in [1, 32]");
    Ctx& c = h->c;
    GPRT_HIP_TRY(hipSetDevice(c.device));
    CtxExtra* x = extra(&c);
    hipStream_t s = c.stream;
    const bool dev = flags & GPRT_PTR_DEVICE;
    c.stats = gprt_hip_run_stats{};
    if (n == 0) return GPRT_OK;
    const float4* d_pm = (const float4*)pos_mass;
    if (!dev) {
        GPRT_TRY(x->q_in.reserve(16ull * n));
        GPRT_HIP_TRY(hipMemcpyAsync(x->q_in.ptr, pos_mass, 16ull * n, hipMemcpyHostToDevice, s));
        d_pm = x->q_in.as<float4>();
    }
    float* d_acc = acc_out;
    if (!dev) {
        GPRT_TRY(x->bh_acc.reserve(12ull * n));
        d_acc = x->bh_acc.as<float>();
    }
    GPRT_TRY(x->bh_xyz.reserve(12ull * n));
    GPRT_TRY(x->flag.reserve(16));
    GPRT_HIP_TRY(hipMemsetAsync(x->flag.ptr, 0, 16, s));
    GPRT_TRY(bh_prepare(&c, d_pm, (int)n, x->bh_xyz.as<float>(), x->flag.as<unsigned>()));
    launch_validate(s, x->bh_xyz.as<float>(), 3 * n, x->flag.as<unsigned>());
    unsigned bad = 0;
    GPRT_HIP_TRY(hipMemcpyAsync(&bad, x->flag.ptr, 4, hipMemcpyDeviceToHost, s));
    GPRT_HIP_TRY(hipStreamSynchronize(s));
    if (bad) return fail(GPRT_ERR_INPUT, "bodies contain NaN/Inf, |coordinate| > 1e18 or a negative mass");
    int64_t off[2] = {0, n};
    GPRT_HIP_TRY(hipEventRecord(c.ev2, s));
    x->bh_tree.build_wide = false;
    GPRT_TRY(tree_build(&c, x->bh_xyz.as<float>(), off, 1, leaf_size, &x->bh_tree));
    GPRT_TRY(bh_run(&c, x->bh_tree, d_pm, (int)n, theta, eps, G, d_acc, x->bh_mass, (flags & GPRT_COLLECT_COUNTERS) != 0,
                          (flags & GPRT_BH_REFERENCE_LAW) != 0, (flags & GPRT_BH_LANE_WALK) != 0));
    GPRT_HIP_TRY(hipEventRecord(c.ev3, s));
    if (!dev) GPRT_HIP_TRY(hipMemcpyAsync(acc_out, d_acc, 12ull * n, hipMemcpyDeviceToHost, s));
    GPRT_TRY(finish_stats(c, n, true));
    float ms = 0;
    if (hipEventElapsedTime(&ms, c.ev2, c.ev3) == hipSuccess) c.stats.ms_total = ms;
    c.stats.ms_build = c.stats.ms_total - c.stats.ms_query;
    return GPRT_OK;
}
