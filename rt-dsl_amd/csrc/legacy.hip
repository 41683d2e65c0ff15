// legacy.hip -- the 13 entry points of the reference's native FFI
// (gprt-optix/src/lib.rs:13-31, implemented in gprt-optix/host/main.cpp:101-415)
// on top of the LBVH and the HIP kernels of this library, so a Rust crate that
// binds them keeps working unchanged.
//
// Device programs.  OptiX-IR cannot be loaded here; `ir_data` instead names one
// of the built-in programs that the reference's macros generate:
//   "gprt-hip:knn-heap"  register-heap k-NN of gprt_rnn_run!
//                        (gprt-macros/src/lib.rs:305-377): per query the k = cap/2
//                        nearest points with d64 < r^2 (strict, like the AnyHit
//                        insert of :357), ids[0..k) then dists[k..2k) as float
//                        bits; empty slots hold 0xFFFFFFFF and fl32(r*r) (:317-320);
//                        lengths = number found (:329-331).  Default program.
//   "gprt-hip:push"      global push of gprt_rnn_run!/dispatch (:276-295): every
//                        point with d64 <= r^2 (the IS test of :341), first `cap`
//                        of them per query in (d64, id) order; lengths = total
//                        count, like the atomicAdd counter.  An array
//                        registered for one query and executed with several
//                        rays holds the hits of all rays in one list (query
//                        order), as gprt_trace! (:637-665) shares one counter.
//   "gprt-hip:autorope"  the Barnes-Hut walk of barnes_hut! (:760-806) over the
//                        caller's flattened octree.
//   "gprt-hip:ir:" + JSON a program of the reference's shader IR, lowered to HIP
//                        and compiled at pipeline creation (codegen.hip); it runs
//                        one lane per query over the LBVH with the arrays
//                        registered on the pipeline, in registration order.
// Every distance uses the binary64 contract d64 of the rest of the library, and
// r^2 = (double)r * (double)r.
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "gprt_internal.hpp"
#include "knn_internal.hpp"

namespace gprt {
int range_run(Ctx* ctx, const Tree& T, const float* d_q, int nq, float r, int64_t* d_off, uint32_t* d_ids,
              int64_t capacity, DevBuf& counts_buf, DevBuf& d2_buf, DevBuf& big_buf, DevBuf& keys_buf,
              DevBuf& perm_buf, DevBuf& qs4_buf, DevBuf& sums_buf, int64_t* h_total, const float* d_radii, bool count = false, bool time_walk = false);
}

using namespace gprt;

enum { PROG_KNN_HEAP = 0, PROG_PUSH = 1, PROG_AUTOROPE = 2, PROG_GENERATED = 3 };

namespace gprt {
int codegen_program(const char* ir, size_t n, std::string& src);
int codegen_compile(const std::string& src, const char* arch, std::vector<char>& code);
}

// host mirror of the generated kernel's launch block (codegen.hip)
struct GenBundle {
    unsigned long long dyn_ptrs[8];
    unsigned int dyn_caps[8];
    unsigned long long dyn_lens[8];
    unsigned long long val_ptrs[8];
};
struct GenLaunch {
    const void* nodes;
    const void* pts;
    int root_ref, root_count;
    const void* rays;
    int num_rays;
    void* bundle;
    int max_steps;
    unsigned* err;
};

struct LegacyArray {
    DevBuf data, lens;
    unsigned cap = 0;
    int nq = 0;
};

struct CGprtPipeline {
    gprt_hip_ctx* ctx = nullptr;
    int program = PROG_KNN_HEAP;
    bool is_hw_triangle = false;
    std::map<std::string, LegacyArray> arrays;
    std::vector<std::string> order;  // registration order (array indices of generated programs)
    hipModule_t module = nullptr;    // generated program
    hipFunction_t kernel = nullptr;
    DevBuf bundle;
    std::map<std::string, DevBuf> values;
    DevBuf q4, xyz, radii, off, tmp_ids, counts, d2, big, keys, perm, qs4, sums;
    DevBuf bodies, spatial, routing, forces;
};

struct CGprtBvh {
    gprt_hip_tree* tree = nullptr;
    float radius = 0.0f;
    int count = 0;
};


namespace gprt {

// radius-bounded exact k-NN in heap layout, one lane per query (input order)
template <int KMAX>
__global__ __launch_bounds__(256) void heap_kernel(const Node* __restrict__ nodes, const float4* __restrict__ pts,
                                                   int2 root, const float4* __restrict__ q4, int nq, int k,
                                                   unsigned* __restrict__ heap, unsigned* __restrict__ lens,
                                                   int max_steps, unsigned* err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const float4 q = q4[i];
    const double r2 = (double)q.w * (double)q.w;
    // slots [0, KMAX-k) are -inf sentinels; the k live slots start at (r2, 0) so
    // only points with d64 < r2 can enter (strict, as the reference's insert)
    double d[KMAX];
    uint32_t id[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
        d[j] = j < KMAX - k ? -INFINITY : r2;
        id[j] = 0u;
    }
    auto insert = [&](double dd, uint32_t ii) {
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
    };
    auto scan = [&](int f, int m) {
        for (int t = 0; t < m; ++t) {
            const float4 p = pts[f + t];
            const double dd = d2d(q.x, q.y, q.z, p.x, p.y, p.z);
            const uint32_t ii = (uint32_t)__float_as_int(p.w);
            if (dd < r2 && (dd < d[KMAX - 1] || (dd == d[KMAX - 1] && ii < id[KMAX - 1]))) insert(dd, ii);
        }
    };
    if (root.y > 0) {
        if (root.x < 0) {
            scan(~root.x, root.y);
        } else {
            int stk[EXACT_STACK];
            int sp = 0, node = root.x, steps = 0;
            while (true) {
                if (++steps > max_steps || sp >= EXACT_STACK) { atomicAdd(err, 1u); break; }
                const Node nd = nodes[node];
                const double wd = d[KMAX - 1];
                const double dA = box2d(q.x, q.y, q.z, nd.alo, nd.ahi);
                const double dB = box2d(q.x, q.y, q.z, nd.blo, nd.bhi);
                const int rA = __float_as_int(nd.alo.w), rB = __float_as_int(nd.blo.w);
                if (rA < 0 && dA <= wd) scan(~rA, __float_as_int(nd.ahi.w));
                if (rB < 0 && dB <= d[KMAX - 1]) scan(~rB, __float_as_int(nd.bhi.w));
                const double w2 = d[KMAX - 1];
                const bool gA = rA >= 0 && dA <= w2, gB = rB >= 0 && dB <= w2;
                if (gA && gB) {
                    const bool nearA = dA <= dB;
                    stk[sp++] = nearA ? rB : rA;
                    node = nearA ? rA : rB;
                    continue;
                }
                if (gA) { node = rA; continue; }
                if (gB) { node = rB; continue; }
                if (sp == 0) break;
                node = stk[--sp];
            }
        }
    }
    const float rsq = q.w * q.w;
    unsigned cnt = 0;
    unsigned* h = heap + (size_t)i * 2 * k;
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
        if (j >= KMAX - k) {
            const int o = j - (KMAX - k);
            const bool real = d[j] < r2;
            h[o] = real ? id[j] : 0xFFFFFFFFu;
            h[k + o] = real ? __float_as_uint((float)d[j]) : __float_as_uint(rsq);
            cnt += real;
        }
    lens[i] = cnt;
}

__global__ void split_queries_kernel(const float4* __restrict__ q4, int nq, float* __restrict__ xyz,
                                     float* __restrict__ radii) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nq) {
        const float4 q = q4[i];
        xyz[3 * i] = q.x;
        xyz[3 * i + 1] = q.y;
        xyz[3 * i + 2] = q.z;
        radii[i] = q.w;
    }
}

// copy the first `cap` ids of every CSR segment into the per-query slot layout
__global__ void push_pack_kernel(const int64_t* __restrict__ off, const uint32_t* __restrict__ ids, int nq,
                                 unsigned cap, unsigned* __restrict__ out, unsigned* __restrict__ lens) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const int64_t b = off[i], m = off[i + 1] - b;
    for (int64_t j = 0; j < m && j < cap; ++j) out[(size_t)i * cap + j] = ids[b + j];
    lens[i] = (unsigned)m;
}

// ---- autorope Barnes-Hut of the reference (gprt-macros/src/lib.rs:760-806) ----
struct __attribute__((packed)) LBody { float px, py, pz, mass, vx, vy, vz; };            // 28 B
struct LSpatial { float cx, cy, cz, mass; };                                             // 16 B
struct LRouting { unsigned next_idx, autorope, is_leaf; float width; unsigned bucket_start, num_particles, pad[2]; };

__global__ void autorope_kernel(const LBody* __restrict__ bodies, int nb, const LSpatial* __restrict__ sp,
                                const LRouting* __restrict__ rt, int nn, float theta, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const LBody me = bodies[i];
    float fx = 0.f, fy = 0.f, fz = 0.f;
    const float G = 6.674e-11f;  // hard-coded in the reference shader (lib.rs:785)
    auto interact = [&](float x, float y, float z, float m) {
        const float dx = x - me.px, dy = y - me.py, dz = z - me.pz;
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > 1e-10f) {
            const float f = (G * me.mass * m) / (d2 * sqrtf(d2 + 1e-6f));
            fx += dx * f;
            fy += dy * f;
            fz += dz * f;
        }
    };
    unsigned node = 0, steps = 0;
    while (node < (unsigned)nn && steps++ < 4u * (unsigned)nn + 8u) {
        const LSpatial s = sp[node];
        const LRouting r = rt[node];
        const float dx = s.cx - me.px, dy = s.cy - me.py, dz = s.cz - me.pz;
        const float d = sqrtf(dx * dx + dy * dy + dz * dz);
        // the reference's hit test: a node of width w is hit by a ray of
        // length max(1e-6, theta d) + 1e-5 when w lies within it
        const bool accept = r.width < fmaxf(1e-6f, theta * d) + 1e-5f;
        if (r.is_leaf == 1) {
            for (unsigned j = 0; j < r.num_particles; ++j) {
                const unsigned b = r.bucket_start + j;
                if (b < (unsigned)nb) interact(bodies[b].px, bodies[b].py, bodies[b].pz, bodies[b].mass);
            }
            node = accept ? r.autorope : r.next_idx;
        } else if (accept) {
            interact(s.cx, s.cy, s.cz, s.mass);
            node = r.autorope;
        } else {
            node = r.next_idx;
        }
    }
    out[3 * i] = fx;
    out[3 * i + 1] = fy;
    out[3 * i + 2] = fz;
}

}  // namespace gprt

static const char kIrPrefix[] = "gprt-hip:ir:";

static int parse_program(const void* ir, size_t n) {
    if (!ir || n == 0) return PROG_KNN_HEAP;
    std::string s((const char*)ir, n);
    while (!s.empty() && (s.back() == '\0' || s.back() == '\n')) s.pop_back();
    if (s == "gprt-hip:push") return PROG_PUSH;
    if (s == "gprt-hip:autorope") return PROG_AUTOROPE;
    if (s.compare(0, sizeof(kIrPrefix) - 1, kIrPrefix) == 0) return PROG_GENERATED;
    return PROG_KNN_HEAP;
}

// lower, compile for the context's GPU and load a generated program
static int load_generated(CGprtPipeline* p, const char* ir, size_t n) {
    std::string src;
    const size_t skip = sizeof(kIrPrefix) - 1;
    GPRT_TRY(codegen_program(ir + skip, n - skip, src));
    hipDeviceProp_t prop;
    GPRT_HIP_TRY(hipGetDeviceProperties(&prop, p->ctx->c.device));
    std::string arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    std::vector<char> code;
    GPRT_TRY(codegen_compile(src, arch.c_str(), code));
    GPRT_HIP_TRY(hipModuleLoadData(&p->module, code.data()));
    GPRT_HIP_TRY(hipModuleGetFunction(&p->kernel, p->module, "gprt_generated"));
    return GPRT_OK;
}

extern "C" CGprtPipeline* gprt_pipeline_create(const void* ir_data, size_t ir_size, int is_hw_triangle) {
    gprt_hip_ctx* ctx = nullptr;
    if (gprt_hip_ctx_create(0, &ctx) != GPRT_OK) return nullptr;
    auto* p = new CGprtPipeline();
    p->ctx = ctx;
    p->program = parse_program(ir_data, ir_size);
    p->is_hw_triangle = is_hw_triangle != 0;
    if (p->program == PROG_GENERATED) {
        size_t n = ir_size;
        while (n && ((const char*)ir_data)[n - 1] == '\0') --n;
        if (load_generated(p, (const char*)ir_data, n) != GPRT_OK) {
            const std::string msg = gprt_hip_last_error();
            gprt_pipeline_destroy(p);
            fail(GPRT_ERR_ARG, msg);
            return nullptr;
        }
    }
    return p;
}

extern "C" void gprt_pipeline_destroy(CGprtPipeline* p) {
    if (!p) return;
    for (auto& kv : p->arrays) {
        kv.second.data.release();
        kv.second.lens.release();
    }
    for (auto& kv : p->values) kv.second.release();
    for (DevBuf* b : {&p->q4, &p->xyz, &p->radii, &p->off, &p->tmp_ids, &p->counts, &p->d2, &p->big, &p->keys,
                      &p->perm, &p->qs4, &p->sums, &p->bodies, &p->spatial, &p->routing, &p->forces, &p->bundle})
        b->release();
    if (p->module) (void)hipModuleUnload(p->module);
    gprt_hip_ctx_destroy(p->ctx);
    delete p;
}

static CGprtBvh* build_from_centres(CGprtPipeline* p, const std::vector<float>& xyz, float radius) {
    auto* b = new CGprtBvh();
    b->radius = radius;
    b->count = (int)(xyz.size() / 3);
    if (gprt_hip_tree_build(p->ctx, xyz.data(), b->count, 8, GPRT_PTR_HOST, &b->tree) != GPRT_OK) {
        delete b;
        return nullptr;
    }
    return b;
}

extern "C" CGprtBvh* gprt_bvh_build(CGprtPipeline* p, const float* h_geom, const float* h_aabbs, int count,
                                    int geom_bytes) {
    if (!p || count < 0) {
        fail(GPRT_ERR_ARG, "gprt_bvh_build: bad pipeline or count");
        return nullptr;
    }
    GPRT_CTX_LOCK(p->ctx);
    std::vector<float> xyz(3 * (size_t)count);
    float radius = 0.0f;
    if (h_geom && geom_bytes >= 16 * count) {
        for (int i = 0; i < count; ++i)
            for (int a = 0; a < 3; ++a) xyz[3 * i + a] = h_geom[4 * i + a];
        if (count) radius = h_geom[3];
    } else if (h_aabbs) {
        for (int i = 0; i < count; ++i)
            for (int a = 0; a < 3; ++a) xyz[3 * i + a] = 0.5f * (h_aabbs[6 * i + a] + h_aabbs[6 * i + 3 + a]);
        if (count) radius = 0.5f * (h_aabbs[3] - h_aabbs[0]);
    } else if (count) {
        fail(GPRT_ERR_ARG, "gprt_bvh_build: no geometry");
        return nullptr;
    }
    return build_from_centres(p, xyz, radius);
}

extern "C" CGprtBvh* gprt_bvh_build_triangles(CGprtPipeline* p, const float* h_verts, const unsigned int* h_indices,
                                              int num_triangles, const float* h_aabbs) {
    (void)h_verts;
    (void)h_indices;
    // the reference forwards to the AABB build (main.cpp:250-252)
    return gprt_bvh_build(p, nullptr, h_aabbs, num_triangles, 0);
}

extern "C" void gprt_bvh_refit(CGprtPipeline* p, CGprtBvh* b, float radius, int count) {
    if (!b) return;
    std::unique_lock<std::recursive_mutex> lk;
    if (p) lk = std::unique_lock<std::recursive_mutex>(p->ctx->mu);
    // the LBVH indexes point centres; a new search radius needs no refit
    b->radius = radius;
    (void)count;
}

extern "C" void gprt_bvh_destroy(CGprtBvh* b) {
    if (!b) return;
    gprt_hip_tree_destroy(b->tree);
    delete b;
}

extern "C" void gprt_register_array(CGprtPipeline* p, const char* name, unsigned int cap, int nq) {
    if (!p || !name || nq < 0) {
        fail(GPRT_ERR_ARG, "gprt_register_array: bad argument");
        return;
    }
    GPRT_CTX_LOCK(p->ctx);
    (void)hipSetDevice(p->ctx->c.device);
    if (!p->arrays.count(name)) p->order.push_back(name);
    LegacyArray& a = p->arrays[name];
    a.cap = cap;
    a.nq = nq;
    if (a.data.reserve(4ull * cap * std::max(nq, 1)) != GPRT_OK || a.lens.reserve(4ull * std::max(nq, 1)) != GPRT_OK)
        return;
    (void)hipMemset(a.data.ptr, 0, 4ull * cap * std::max(nq, 1));
    (void)hipMemset(a.lens.ptr, 0, 4ull * std::max(nq, 1));
}

extern "C" void gprt_register_value(CGprtPipeline* p, const char* name) {
    if (!p || !name) return;
    GPRT_CTX_LOCK(p->ctx);
    DevBuf& v = p->values[name];
    if (v.reserve(4) == GPRT_OK) (void)hipMemset(v.ptr, 0, 4);
}

extern "C" void gprt_execute(CGprtPipeline* p, CGprtBvh* b, const float* h_queries, int count) {
    if (!p || !b || count < 0 || (count && !h_queries)) {
        fail(GPRT_ERR_ARG, "gprt_execute: bad argument");
        return;
    }
    GPRT_CTX_LOCK(p->ctx);
    if (p->program == PROG_AUTOROPE) {
        fail(GPRT_ERR_STATE, "gprt_execute: the autorope program runs through gprt_execute_autorope_soa_ffi");
        return;
    }
    Ctx& c = p->ctx->c;
    (void)hipSetDevice(c.device);
    hipStream_t s = c.stream;
    if (p->q4.reserve(16ull * std::max(count, 1)) != GPRT_OK) return;
    if (count && hipMemcpyAsync(p->q4.ptr, h_queries, 16ull * count, hipMemcpyHostToDevice, s) != hipSuccess) {
        fail(GPRT_ERR_HIP, "gprt_execute: query upload failed");
        return;
    }
    const Tree& T = b->tree->t;
    unsigned* counters = c.counters.as<unsigned>();
    (void)launch_fill_u32(s, counters, CNT_TOTAL + 2 * CNT_STATS64, 0u);
    int2 root;
    (void)hipMemcpyAsync(&root, T.roots.ptr, sizeof(int2), hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    if (p->program == PROG_GENERATED) {
        if (p->order.size() > 8) {
            fail(GPRT_ERR_ARG, "gprt_execute: a generated program binds at most 8 arrays");
            return;
        }
        GenBundle hb;
        memset(&hb, 0, sizeof(hb));
        for (size_t j = 0; j < p->order.size(); ++j) {
            LegacyArray& a = p->arrays[p->order[j]];
            if (a.nq < count) {
                fail(GPRT_ERR_ARG, "gprt_execute: array " + p->order[j] + " holds fewer queries than the launch");
                return;
            }
            hb.dyn_ptrs[j] = (unsigned long long)(uintptr_t)a.data.ptr;
            hb.dyn_caps[j] = a.cap;
            hb.dyn_lens[j] = (unsigned long long)(uintptr_t)a.lens.ptr;
            (void)launch_fill_u32(s, a.lens.as<unsigned>(), count, 0u);
        }
        if (p->bundle.reserve(sizeof(GenBundle)) != GPRT_OK) return;
        if (hipMemcpyAsync(p->bundle.ptr, &hb, sizeof(hb), hipMemcpyHostToDevice, s) != hipSuccess) {
            fail(GPRT_ERR_HIP, "gprt_execute: bundle upload failed");
            return;
        }
        GenLaunch L{T.nodes.ptr, T.pts4.ptr, root.x, root.y, p->q4.ptr, count, p->bundle.ptr, T.n + 64,
                    counters + CNT_ERROR};
        void* args[] = {&L};
        if (count && hipModuleLaunchKernel(p->kernel, (count + 255) / 256, 1, 1, 256, 1, 1, 0, s, args, nullptr) !=
                         hipSuccess) {
            fail(GPRT_ERR_HIP, "gprt_execute: generated program launch failed");
            return;
        }
    }
    for (auto& kv : p->arrays) {
        LegacyArray& a = kv.second;
        // push program, one-query array, several rays: the hits of all rays
        // share one list (gprt_trace!, gprt-macros/src/lib.rs:637-665)
        const bool flat = p->program == PROG_PUSH && a.nq == 1 && count > 1;
        if (p->program == PROG_GENERATED || (a.nq < count && !flat) || count == 0) continue;
        if (p->program == PROG_KNN_HEAP) {
            const int k = (int)(a.cap / 2);
            if (k < 1 || k > GPRT_HIP_K_MAX) {
                fail(GPRT_ERR_ARG, "gprt_execute: register-heap capacity must be 2k with 1 <= k <= " +
                                       std::to_string(GPRT_HIP_K_MAX));
                return;
            }
            if (k > 64) {  // one query per wavefront, the list spread over its lanes
                if (launch_heap_wave(s, T.nodes.as<Node>(), T.pts4.as<float4>(), root, p->q4.as<float4>(), count, k,
                                     a.data.as<unsigned>(), a.lens.as<unsigned>(), T.n + 64,
                                     counters + CNT_ERROR) != GPRT_OK)
                    return;
                continue;
            }
            dim3 grid((count + 255) / 256);
#define HEAPK(KM)                                                                                              \
    hipLaunchKernelGGL(heap_kernel<KM>, grid, dim3(256), 0, s, T.nodes.as<Node>(), T.pts4.as<float4>(), root, \
                       p->q4.as<float4>(), count, k, a.data.as<unsigned>(), a.lens.as<unsigned>(), T.n + 64,  \
                       counters + CNT_ERROR)
            if (k <= 8) HEAPK(8);
            else if (k <= 16) HEAPK(16);
            else if (k <= 32) HEAPK(32);
            else HEAPK(64);
#undef HEAPK
        } else {
            if (p->xyz.reserve(12ull * count) != GPRT_OK || p->radii.reserve(4ull * count) != GPRT_OK ||
                p->off.reserve(8ull * (count + 1)) != GPRT_OK)
                return;
            hipLaunchKernelGGL(split_queries_kernel, dim3((count + 255) / 256), dim3(256), 0, s, p->q4.as<float4>(),
                               count, p->xyz.as<float>(), p->radii.as<float>());
            int64_t total = 0;
            if (range_run(&c, T, p->xyz.as<float>(), count, 0.0f, p->off.as<int64_t>(), nullptr, 0, p->counts, p->d2,
                          p->big, p->keys, p->perm, p->qs4, p->sums, &total, p->radii.as<float>()) != GPRT_OK)
                return;
            if (p->tmp_ids.reserve(4ull * std::max<int64_t>(total, 1)) != GPRT_OK) return;
            if (range_run(&c, T, p->xyz.as<float>(), count, 0.0f, p->off.as<int64_t>(), p->tmp_ids.as<uint32_t>(),
                          total, p->counts, p->d2, p->big, p->keys, p->perm, p->qs4, p->sums, nullptr,
                          p->radii.as<float>()) != GPRT_OK)
                return;
            if (flat) {
                // query-major order, each ray's hits by (d64, id); the length
                // counts every hit, the list keeps the first cap
                const int64_t m = std::min<int64_t>(total, (int64_t)a.cap);
                if (m > 0 && hipMemcpyAsync(a.data.ptr, p->tmp_ids.ptr, 4ull * m, hipMemcpyDeviceToDevice, s) !=
                                 hipSuccess) {
                    fail(GPRT_ERR_HIP, "gprt_execute: copy failed");
                    return;
                }
                (void)launch_fill_u32(s, a.lens.as<unsigned>(), 1, (unsigned)std::min<int64_t>(total, 0xFFFFFFFFll));
            } else {
                hipLaunchKernelGGL(push_pack_kernel, dim3((count + 255) / 256), dim3(256), 0, s,
                                   p->off.as<int64_t>(), p->tmp_ids.as<uint32_t>(), count, a.cap,
                                   a.data.as<unsigned>(), a.lens.as<unsigned>());
            }
        }
    }
    unsigned err = 0;
    (void)hipMemcpyAsync(&err, counters + CNT_ERROR, 4, hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) fail(GPRT_ERR_HIP, "gprt_execute: kernel failed");
    if (err) fail(GPRT_ERR_STATE, "gprt_execute: traversal step guard tripped");
}

extern "C" void gprt_retrieve_array_lengths(CGprtPipeline* p, const char* name, unsigned int* h_lengths,
                                            int num_queries) {
    if (!p || !name || !h_lengths) return;
    GPRT_CTX_LOCK(p->ctx);
    auto it = p->arrays.find(name);
    if (it == p->arrays.end()) {
        fail(GPRT_ERR_ARG, std::string("unknown array ") + name);
        return;
    }
    int m = std::min(num_queries, it->second.nq);
    if (m > 0) (void)hipMemcpy(h_lengths, it->second.lens.ptr, 4ull * m, hipMemcpyDeviceToHost);
}

extern "C" void gprt_retrieve_array_flat(CGprtPipeline* p, const char* name, unsigned int* h_out, int total) {
    if (!p || !name || !h_out) return;
    GPRT_CTX_LOCK(p->ctx);
    auto it = p->arrays.find(name);
    if (it == p->arrays.end()) {
        fail(GPRT_ERR_ARG, std::string("unknown array ") + name);
        return;
    }
    size_t m = std::min<size_t>((size_t)std::max(total, 0), (size_t)it->second.cap * it->second.nq);
    if (m) (void)hipMemcpy(h_out, it->second.data.ptr, 4 * m, hipMemcpyDeviceToHost);
}

extern "C" void gprt_retrieve_value(CGprtPipeline* p, const char* name, unsigned int* h_out) {
    if (!p || !name || !h_out) return;
    GPRT_CTX_LOCK(p->ctx);
    auto it = p->values.find(name);
    if (it == p->values.end()) {
        fail(GPRT_ERR_ARG, std::string("unknown value ") + name);
        return;
    }
    (void)hipMemcpy(h_out, it->second.ptr, 4, hipMemcpyDeviceToHost);
}

extern "C" void gprt_execute_autorope_soa_ffi(CGprtPipeline* p, const void* h_bodies, int nb, const void* h_spatial,
                                              int nn, const void* h_routing, float theta, void* h_out) {
    if (!p || nb < 0 || nn < 0 || (nb && (!h_bodies || !h_out)) || (nn && (!h_spatial || !h_routing))) {
        fail(GPRT_ERR_ARG, "gprt_execute_autorope_soa_ffi: bad argument");
        return;
    }
    GPRT_CTX_LOCK(p->ctx);
    Ctx& c = p->ctx->c;
    (void)hipSetDevice(c.device);
    hipStream_t s = c.stream;
    if (nb == 0) return;
    if (p->bodies.reserve(28ull * nb) != GPRT_OK || p->spatial.reserve(16ull * std::max(nn, 1)) != GPRT_OK ||
        p->routing.reserve(32ull * std::max(nn, 1)) != GPRT_OK || p->forces.reserve(12ull * nb) != GPRT_OK)
        return;
    (void)hipMemcpyAsync(p->bodies.ptr, h_bodies, 28ull * nb, hipMemcpyHostToDevice, s);
    if (nn) {
        (void)hipMemcpyAsync(p->spatial.ptr, h_spatial, 16ull * nn, hipMemcpyHostToDevice, s);
        (void)hipMemcpyAsync(p->routing.ptr, h_routing, 32ull * nn, hipMemcpyHostToDevice, s);
    }
    hipLaunchKernelGGL(autorope_kernel, dim3((nb + 255) / 256), dim3(256), 0, s, (const LBody*)p->bodies.ptr, nb,
                       (const LSpatial*)p->spatial.ptr, (const LRouting*)p->routing.ptr, nn, theta,
                       p->forces.as<float>());
    (void)hipMemcpyAsync(h_out, p->forces.ptr, 12ull * nb, hipMemcpyDeviceToHost, s);
    if (hipStreamSynchronize(s) != hipSuccess) fail(GPRT_ERR_HIP, "autorope kernel failed");
}
