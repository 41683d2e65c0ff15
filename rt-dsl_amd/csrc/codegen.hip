// codegen.hip -- device programs generated from the reference's shader IR and
// compiled at run time with hiprtc for the GPU of the context.
//
// The reference lowers a tree of ShaderNode values (gprt-ir/src/lib.rs:8-18:
// TraceRay, TerminateRay, IgnoreIntersection, PushToDynamicArray,
// UpdatePayload, If, Block, RawCuda) to a CUDA translation unit of OptiX
// programs (gprt-codegen/src/lib.rs:78-119, 121-182) and runs nvcc on it
// (:184-203).  Here the same tree, serialised as JSON, is lowered to HIP
// statements placed inside one generated traversal kernel over the LBVH:
//
//   raygen        runs once per query (one lane per query); TraceRay inside it
//                 is the walk, followed by its on_hit / on_miss branches
//   intersection  runs for every point of every leaf bucket the walk enters
//                 (the walk enters a box when its FP32 distance to the query is
//                 within the bound in payload register p2, as the OptiX box test
//                 with a search radius); ReportIntersection accepts the point
//   anyhit        runs for every accepted point
//   closesthit / miss run after the walk when some / no point was accepted
//
// Names visible to program code: gprt_qid (query index), gprt_ray (float4:
// query xyz and radius), p0..p3 (unsigned payload registers: p1 starts at
// 0xFFFFFFFF, p2 at the bits of radius^2, the walk's bound), bundle (the
// registered arrays: dyn_ptrs / dyn_caps / dyn_lens, in registration order),
// and in intersection / anyhit gprt_prim (point id), gprt_sphere (float4
// point, id bits in w) and gprt_dist_sq ((dx*dx + dy*dy) + dz*dz in FP32).
//
// JSON forms of the nodes: "IgnoreIntersection", "TerminateRay",
// "ReportIntersection", {"Block": [node...]},
// {"If": {"condition": "c", "then": node, "else": node}},
// {"PushToDynamicArray": {"array": i, "value": "v"}},
// {"UpdatePayload": {"register": r, "value": "v"}}, {"Raw": "statements"},
// {"TraceRay": {"on_hit": node, "on_miss": node}}.  A program is
// {"raygen": node, "intersection": node, "anyhit": node, "closesthit": node,
//  "miss": node} with every entry but raygen optional.
#include <hip/hiprtc.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "gprt_internal.hpp"
#include "json.hpp"

namespace gprt {
namespace {

// ---------------------------------------------------------------------------
// lowering of ShaderNode trees to HIP statements
// ---------------------------------------------------------------------------
enum Stage { RAYGEN, HIT, AFTER };  // raygen, intersection / anyhit, closesthit / miss

struct Lowerer {
    std::string out;
    std::string err;
    int indent = 1;
    const JVal* prog = nullptr;
    int traces = 0;

    void emit(const std::string& line) {
        out.append(4 * indent, ' ');
        out += line;
        out += '\n';
    }
    bool bad(const std::string& what) {
        if (err.empty()) err = what;
        return false;
    }
    static bool int_of(const JVal* v, int lo, int hi, int& out) {
        if (!v || v->kind != JVal::Num || v->num != (double)(int)v->num) return false;
        out = (int)v->num;
        return out >= lo && out <= hi;
    }
    bool text_of(const JVal* v, const char* what, std::string& out) {
        if (!v || v->kind != JVal::Str || v->str.empty()) return bad(std::string(what) + " must be a non-empty string");
        out = v->str;
        return true;
    }
    bool block(const JVal* v, Stage st) {
        emit("{");
        ++indent;
        bool ok = v ? node(*v, st) : true;
        --indent;
        emit("}");
        return ok;
    }
    bool node(const JVal& v, Stage st) {
        if (v.kind == JVal::Null) return true;
        if (v.kind == JVal::Str) {
            if (v.str == "IgnoreIntersection") {
                if (st != HIT) return bad("IgnoreIntersection outside intersection / anyhit");
                emit("gprt_ignore = true;");
                return true;
            }
            if (v.str == "ReportIntersection") {
                if (st != HIT) return bad("ReportIntersection outside intersection / anyhit");
                emit("gprt_reported = true;");
                return true;
            }
            if (v.str == "TerminateRay") {
                if (st != HIT) return bad("TerminateRay outside intersection / anyhit");
                emit("gprt_terminate = true;");
                return true;
            }
            return bad("unknown node \"" + v.str + "\"");
        }
        if (v.kind != JVal::Obj || v.obj.size() != 1) return bad("a node is a string or a one-key object");
        const std::string& kind = v.obj[0].first;
        const JVal& a = v.obj[0].second;
        if (kind == "Block") {
            if (a.kind != JVal::Arr) return bad("Block takes an array");
            for (auto& item : a.arr)
                if (!node(item, st)) return false;
            return true;
        }
        if (kind == "Raw" || kind == "RawCuda") {
            if (a.kind != JVal::Str) return bad("Raw takes a string");
            size_t b = 0;
            while (b <= a.str.size()) {
                size_t e = a.str.find('\n', b);
                if (e == std::string::npos) e = a.str.size();
                emit(a.str.substr(b, e - b));
                b = e + 1;
            }
            return true;
        }
        if (kind == "If") {
            std::string cond;
            if (a.kind != JVal::Obj || !text_of(a.get("condition"), "If.condition", cond)) return false;
            emit("if (" + cond + ")");
            if (!block(a.get("then"), st)) return false;
            if (a.get("else")) {
                emit("else");
                if (!block(a.get("else"), st)) return false;
            }
            return true;
        }
        if (kind == "PushToDynamicArray") {
            int arr = 0;
            std::string val;
            if (a.kind != JVal::Obj || !int_of(a.get("array"), 0, 7, arr))
                return bad("PushToDynamicArray.array must be an integer in [0, 8)");
            if (!text_of(a.get("value"), "PushToDynamicArray.value", val)) return false;
            const std::string A = std::to_string(arr);
            emit("{");
            emit("    unsigned int* gprt_len = (unsigned int*)bundle->dyn_lens[" + A + "] + gprt_qid;");
            emit("    const unsigned int gprt_at = (*gprt_len)++;");
            emit("    if (gprt_at < bundle->dyn_caps[" + A + "])");
            emit("        ((unsigned int*)bundle->dyn_ptrs[" + A + "])[(size_t)gprt_qid * bundle->dyn_caps[" + A +
                 "] + gprt_at] = (unsigned int)(" + val + ");");
            emit("}");
            return true;
        }
        if (kind == "UpdatePayload") {
            int reg = 0;
            std::string val;
            if (a.kind != JVal::Obj || !int_of(a.get("register"), 0, 3, reg))
                return bad("UpdatePayload.register must be an integer in [0, 4)");
            if (!text_of(a.get("value"), "UpdatePayload.value", val)) return false;
            emit("p" + std::to_string(reg) + " = (unsigned int)(" + val + ");");
            return true;
        }
        if (kind == "TraceRay") {
            if (st != RAYGEN) return bad("TraceRay outside raygen");
            if (a.kind != JVal::Obj) return bad("TraceRay takes an object");
            if (++traces > 4) return bad("at most 4 TraceRay nodes");
            return trace(a.get("on_hit"), a.get("on_miss"));
        }
        return bad("unknown node \"" + kind + "\"");
    }
    // the walk: every point of every leaf bucket whose box lies within the
    // bound in p2 goes through the intersection program
    bool trace(const JVal* on_hit, const JVal* on_miss) {
        emit("{");
        ++indent;
        emit("unsigned int gprt_hits = 0;");
        emit("bool gprt_stop = false;");
        emit("auto gprt_leaf = [&](int gprt_first, int gprt_count) {");
        ++indent;
        emit("for (int gprt_t = 0; gprt_t < gprt_count && !gprt_stop; ++gprt_t) {");
        ++indent;
        emit("const float4 gprt_sphere = params.pts[gprt_first + gprt_t];");
        emit("const unsigned int gprt_prim = (unsigned int)__float_as_int(gprt_sphere.w);");
        emit("const float gprt_dist_sq = gprt_d2(gprt_ray.x, gprt_ray.y, gprt_ray.z, gprt_sphere.x, gprt_sphere.y, "
             "gprt_sphere.z);");
        emit("bool gprt_reported = false, gprt_ignore = false, gprt_terminate = false;");
        emit("(void)gprt_prim; (void)gprt_dist_sq; (void)gprt_ignore;");
        if (!block(prog->get("intersection"), HIT)) return false;
        emit("if (gprt_reported && !gprt_terminate) {");
        ++indent;
        emit("++gprt_hits;");
        if (!block(prog->get("anyhit"), HIT)) return false;
        --indent;
        emit("}");
        emit("if (gprt_terminate) gprt_stop = true;");
        --indent;
        emit("}");
        --indent;
        emit("};");
        emit("gprt_walk(params, gprt_ray, p2, gprt_stop, gprt_leaf);");
        emit("if (gprt_hits) {");
        ++indent;
        if (!block(prog->get("closesthit"), AFTER)) return false;
        if (!block(on_hit, RAYGEN)) return false;
        --indent;
        emit("} else {");
        ++indent;
        if (!block(prog->get("miss"), AFTER)) return false;
        if (!block(on_miss, RAYGEN)) return false;
        --indent;
        emit("}");
        --indent;
        emit("}");
        return true;
    }
};

// Everything the generated program needs besides its own bodies.  The walk
// is the per-lane closest-child-first stack walk of the library's other
// kernels, with the bound re-read from p2 at every step so a program that
// lowers p2 (a register heap) prunes as it goes.
const char* kPrelude = R"HIP(
struct GprtNode { float4 alo, ahi, blo, bhi; };
struct PayloadBundle {
    unsigned long long dyn_ptrs[8];
    unsigned int dyn_caps[8];
    unsigned long long dyn_lens[8];
    unsigned long long val_ptrs[8];
};
struct GprtLaunch {
    const GprtNode* nodes;
    const float4* pts;
    int root_ref, root_count;
    const float4* rays;
    int num_rays;
    PayloadBundle* bundle;
    int max_steps;
    unsigned int* err;
};
__device__ __forceinline__ float gprt_d2(float ax, float ay, float az, float bx, float by, float bz) {
    const float dx = ax - bx, dy = ay - by, dz = az - bz;
    return (dx * dx + dy * dy) + dz * dz;
}
__device__ __forceinline__ float gprt_box(float4 q, float4 lo, float4 hi) {
    return gprt_d2(q.x, q.y, q.z, __builtin_amdgcn_fmed3f(q.x, lo.x, hi.x), __builtin_amdgcn_fmed3f(q.y, lo.y, hi.y),
                   __builtin_amdgcn_fmed3f(q.z, lo.z, hi.z));
}
template <class Leaf>
__device__ __forceinline__ void gprt_walk(const GprtLaunch& params, float4 q, const unsigned int& bound,
                                          const bool& stop, Leaf&& leaf) {
    if (params.root_count <= 0) return;
    if (params.root_ref < 0) {
        leaf(~params.root_ref, params.root_count);
        return;
    }
    int stk[96];
    int sp = 0, node = params.root_ref, steps = 0;
    while (!stop) {
        if (++steps > params.max_steps || sp >= 96) {
            atomicAdd(params.err, 1u);
            return;
        }
        const GprtNode nd = params.nodes[node];
        const float dA = gprt_box(q, nd.alo, nd.ahi), dB = gprt_box(q, nd.blo, nd.bhi);
        const int rA = __float_as_int(nd.alo.w), rB = __float_as_int(nd.blo.w);
        if (rA < 0 && dA <= __uint_as_float(bound)) leaf(~rA, __float_as_int(nd.ahi.w));
        if (rB < 0 && dB <= __uint_as_float(bound) && !stop) leaf(~rB, __float_as_int(nd.bhi.w));
        const float w = __uint_as_float(bound);
        const bool gA = rA >= 0 && dA <= w, gB = rB >= 0 && dB <= w;
        if (gA && gB) {
            const bool nearA = dA <= dB;
            stk[sp++] = nearA ? rB : rA;
            node = nearA ? rA : rB;
            continue;
        }
        if (gA) { node = rA; continue; }
        if (gB) { node = rB; continue; }
        if (sp == 0) return;
        node = stk[--sp];
    }
}
)HIP";

}  // namespace

// IR text -> complete HIP translation unit
int codegen_program(const char* ir, size_t n, std::string& src) {
    JParser jp(ir, n);
    JVal prog;
    if (!jp.value(prog)) return fail(GPRT_ERR_ARG, "codegen: " + jp.err);
    jp.ws();
    if (jp.i != n) return fail(GPRT_ERR_ARG, "codegen: trailing bytes after the program");
    if (prog.kind != JVal::Obj) return fail(GPRT_ERR_ARG, "codegen: the program is a JSON object");
    for (auto& kv : prog.obj)
        if (kv.first != "raygen" && kv.first != "intersection" && kv.first != "anyhit" && kv.first != "closesthit" &&
            kv.first != "miss")
            return fail(GPRT_ERR_ARG, "codegen: unknown program \"" + kv.first + "\"");
    if (!prog.get("raygen")) return fail(GPRT_ERR_ARG, "codegen: a raygen program is required");
    Lowerer lw;
    lw.prog = &prog;
    if (!lw.node(*prog.get("raygen"), RAYGEN)) return fail(GPRT_ERR_ARG, "codegen: " + lw.err);
    src = kPrelude;
    src += "extern \"C\" __global__ __launch_bounds__(256) void gprt_generated(GprtLaunch params) {\n";
    src += "    const unsigned int gprt_qid = blockIdx.x * blockDim.x + threadIdx.x;\n";
    src += "    if (gprt_qid >= (unsigned int)params.num_rays) return;\n";
    src += "    PayloadBundle* bundle = params.bundle;\n";
    src += "    const float4 gprt_ray = params.rays[gprt_qid];\n";
    src += "    unsigned int p0 = 0u, p1 = 0xFFFFFFFFu, p2 = __float_as_uint(gprt_ray.w * gprt_ray.w), p3 = 0u;\n";
    src += "    (void)bundle; (void)p0; (void)p1; (void)p3;\n";
    src += lw.out;
    src += "}\n";
    return GPRT_OK;
}

// HIP translation unit -> code object for `arch`, with the library's own
// floating-point flags (no contraction, no packed FP32)
int codegen_compile(const std::string& src, const char* arch, std::vector<char>& code) {
    hiprtcProgram prog = nullptr;
    if (hiprtcCreateProgram(&prog, src.c_str(), "gprt_generated.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
        return fail(GPRT_ERR_HIP, "hiprtc: cannot create the program");
    const std::string off = std::string("--offload-arch=") + arch;
    const char* opts[] = {off.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
                          "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (rc != HIPRTC_SUCCESS) {
        size_t ls = 0;
        std::string log;
        if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls) {
            log.resize(ls);
            (void)hiprtcGetProgramLog(prog, &log[0]);
        }
        (void)hiprtcDestroyProgram(&prog);
        return fail(GPRT_ERR_ARG, "hiprtc: compilation failed: " + log);
    }
    size_t sz = 0;
    if (hiprtcGetCodeSize(prog, &sz) != HIPRTC_SUCCESS || sz == 0) {
        (void)hiprtcDestroyProgram(&prog);
        return fail(GPRT_ERR_HIP, "hiprtc: no code produced");
    }
    code.resize(sz);
    const bool ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
    (void)hiprtcDestroyProgram(&prog);
    return ok ? GPRT_OK : fail(GPRT_ERR_HIP, "hiprtc: cannot read the code");
}

}  // namespace gprt

using namespace gprt;

extern "C" int gprt_hip_codegen(const char* ir_json, size_t ir_size, char* src_out, size_t src_capacity,
                                size_t* src_size) {
    if (!ir_json || !src_size) return fail(GPRT_ERR_ARG, "null argument");
    std::string src;
    GPRT_TRY(codegen_program(ir_json, ir_size, src));
    *src_size = src.size() + 1;
    if (src_out) {
        if (src_capacity < src.size() + 1) return fail(GPRT_ERR_CAPACITY, "source buffer too small");
        memcpy(src_out, src.c_str(), src.size() + 1);
    }
    return GPRT_OK;
}

extern "C" int gprt_hip_codegen_compile(const char* ir_json, size_t ir_size, const char* arch, size_t* code_size) {
    if (!ir_json || !code_size) return fail(GPRT_ERR_ARG, "null argument");
    std::string src;
    GPRT_TRY(codegen_program(ir_json, ir_size, src));
    std::vector<char> code;
    GPRT_TRY(codegen_compile(src, arch ? arch : "gfx950", code));
    *code_size = code.size();
    return GPRT_OK;
}
