// autotune.hip -- picks the batched k-NN schedule for a data set.
//
// Replaces the reference's hill-climb macro gprt_autotune!
// (gprt-macros/src/lib.rs:120-214: perturb percentile and multiplier, keep a
// change when every query got k ids and the run was faster) and the funnel
// profiler (gprt-profiler/src/main.rs:37-133: grid over percentile x multiplier
// x cap on growing subsets, saturation gate).  Here the knobs are the ones that
// matter on the GPU -- traversal mapping, leaf size, LDS stack ring size, round structure, initial-radius rule
// (percentile / unbounded) and growth multiplier -- and the gate is stronger:
// every candidate's ids must equal those of the reference schedule bit for bit.
// Timing uses hipEvents around the query phase (median of the timed repeats).
// The result is written as a JSON superset of the reference's Schedule /
// wisdom format (gprt-ir/src/lib.rs:58-89).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "gprt_internal.hpp"
#include "json.hpp"

using namespace gprt;

namespace {

struct Cand {
    gprt_hip_config cfg;
    int leaf;
    int sah;  // 1: the SAH tree (GPRT_BUILD_SAH), 0: the LBVH
};

// One tuning problem: builds the tree of a candidate (leaf size, SAH or not)
// and runs one k-NN call with a config into d_ids (nout ids).
struct Problem {
    std::function<int(int leaf, int sah, gprt_hip_tree** out)> build;
    std::function<int(gprt_hip_tree* t, const gprt_hip_config& cfg, uint32_t* d_ids)> run;
    size_t nout = 0;
};

// median query-phase time of three calls after a warm-up
int time_config(gprt_hip_ctx* h, const Problem& P, gprt_hip_tree* t, const gprt_hip_config& cfg, uint32_t* d_ids,
                double* ms_out) {
    std::vector<double> ts;
    for (int r = 0; r < 4; ++r) {
        GPRT_TRY(P.run(t, cfg, d_ids));
        gprt_hip_run_stats st;
        gprt_hip_stats(h, &st);
        if (r > 0) ts.push_back(st.ms_query);
    }
    std::sort(ts.begin(), ts.end());
    *ms_out = ts[ts.size() / 2];
    return GPRT_OK;
}

std::string schedule_json(const gprt_hip_config& c, int leaf, int sah, double ms) {
    char heur[96];
    if (c.radius_heuristic == GPRT_RADIUS_PERCENTILE) snprintf(heur, sizeof heur, "{\"SampledPercentile\": %.6g}", c.radius_param);
    else if (c.radius_heuristic == GPRT_RADIUS_FIXED) snprintf(heur, sizeof heur, "{\"Fixed\": %.9g}", c.radius_param);
    else snprintf(heur, sizeof heur, "\"SampledMax\"");
    char buf[1024];
    snprintf(buf, sizeof buf,
             "{\"radius_increment_mult\": %.6g, \"max_hits_per_query\": 2000, \"use_morton_lbv\": %s, "
             "\"radius_heuristic\": %s, \"memory_strategy\": \"PayloadRegisterHeap\", "
             "\"build_strategy\": \"%s\", \"geom_type\": \"Spheres\", "
             "\"gprt_hip\": {\"sampler\": %d, \"radius_heuristic\": %d, \"radius_param\": %.6g, "
             "\"radius_mult\": %.6g, \"max_rounds\": %d, \"traversal\": %d, \"stack_depth\": %d, "
             "\"guard_slots\": %d, \"sort_queries\": %d, \"use_graph\": %d, \"grid_blocks\": %d, \"lds_stack\": %d, \"compact_rounds\": %d, "
             "\"leaf_size\": %d, \"ms_query\": %.6g}}\n",
             c.radius_mult, c.sort_queries ? "true" : "false", heur, sah ? "PreferFastTrace" : "PreferFastBuild",
             c.sampler, c.radius_heuristic, c.radius_param,
             c.radius_mult, c.max_rounds, c.traversal, c.stack_depth, c.guard_slots, c.sort_queries, c.use_graph,
             c.grid_blocks, c.lds_stack, c.compact_rounds, leaf, ms);
    return buf;
}


// Coordinate descent from the default schedule on the leaf-8 LBVH, one knob at
// a time (like the reference's hill climb), bounded by budget_trials timed
// candidates; every candidate's ids must equal the default's.
int autotune_core(gprt_hip_ctx* h, const Problem& P, int budget_trials, bool with_sah, Cand* best_out,
                  double* best_ms_out) {
    DevBuf dids, dref;
    GPRT_TRY(dids.reserve(4 * std::max<size_t>(P.nout, 1)));
    GPRT_TRY(dref.reserve(4 * std::max<size_t>(P.nout, 1)));
    const int leaves[3] = {4, 8, 16};
    // trees[s * 3 + i]: leaf size leaves[i], s = 1 the SAH tree
    gprt_hip_tree* trees[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int rc = GPRT_OK;
    for (int i = 0; i < (with_sah ? 6 : 3) && rc == GPRT_OK; ++i) rc = P.build(leaves[i % 3], i >= 3, &trees[i]);
    // reference answer: default schedule on the leaf-8 tree
    gprt_hip_config base;
    gprt_hip_config_default(&base);
    if (rc == GPRT_OK) rc = P.run(trees[1], base, dref.as<uint32_t>());
    std::vector<uint32_t> ref, got;
    if (rc == GPRT_OK) {
        ref.resize(P.nout);
        if (hipMemcpy(ref.data(), dref.ptr, 4 * P.nout, hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(GPRT_ERR_HIP, "copy failed");
    }
    Cand best{base, 8, 0};
    double best_ms = 1e30;
    int trials = 0;
    auto tree_of = [&](int leaf, int sah) { return trees[3 * sah + (leaf == 4 ? 0 : (leaf == 8 ? 1 : 2))]; };
    auto try_cand = [&](const Cand& cd) -> int {
        if (trials >= std::max(budget_trials, 1)) return GPRT_OK;
        ++trials;
        double ms = 0;
        GPRT_TRY(time_config(h, P, tree_of(cd.leaf, cd.sah), cd.cfg, dids.as<uint32_t>(), &ms));
        got.resize(P.nout);
        GPRT_HIP_TRY(hipMemcpy(got.data(), dids.ptr, 4 * P.nout, hipMemcpyDeviceToHost));
        if (got != ref) return fail(GPRT_ERR_STATE, "autotune: a candidate schedule changed the result");
        if (ms < best_ms) {
            best_ms = ms;
            best = cd;
        }
        return GPRT_OK;
    };
    if (rc == GPRT_OK) rc = try_cand(best);
    for (int pass = 0; pass < 2 && rc == GPRT_OK; ++pass) {
        const Cand start = best;
        if (with_sah && rc == GPRT_OK) {
            // the tree: SAH splits (slower build, fewer node visits) or the LBVH
            Cand cd = best;
            cd.sah = best.sah ? 0 : 1;
            rc = try_cand(cd);
        }
        for (int leaf : leaves) {
            if (leaf == start.leaf || rc != GPRT_OK) continue;
            Cand cd = best;
            cd.leaf = leaf;
            rc = try_cand(cd);
        }
        for (int ring : {8, 16, 32}) {
            if (ring == best.cfg.lds_stack || rc != GPRT_OK) continue;
            Cand cd = best;
            cd.cfg.lds_stack = ring;
            rc = try_cand(cd);
        }
        if (rc == GPRT_OK) {
            Cand cd = best;
            cd.cfg.compact_rounds = best.cfg.compact_rounds ? 0 : 1;
            rc = try_cand(cd);
        }
        for (int trav : {GPRT_TRAV_THREAD, GPRT_TRAV_WAVE, GPRT_TRAV_WIDE}) {
            if (trav == best.cfg.traversal || rc != GPRT_OK) continue;
            Cand cd = best;
            cd.cfg.traversal = trav;
            rc = try_cand(cd);
        }
        for (float p : {0.5f, 0.8f, 0.95f, 0.99f}) {
            if (rc != GPRT_OK) break;
            Cand cd = best;
            cd.cfg.radius_heuristic = GPRT_RADIUS_PERCENTILE;
            cd.cfg.radius_param = p;
            rc = try_cand(cd);
        }
        if (rc == GPRT_OK) {
            Cand cd = best;
            cd.cfg.radius_heuristic = GPRT_RADIUS_INFINITE;
            rc = try_cand(cd);
        }
        for (float m : {1.5f, 2.0f, 3.0f, 5.0f}) {
            if (rc != GPRT_OK || best.cfg.radius_heuristic == GPRT_RADIUS_INFINITE) break;
            Cand cd = best;
            cd.cfg.radius_mult = m;
            rc = try_cand(cd);
        }
    }
    for (auto* t : trees) gprt_hip_tree_destroy(t);
    dids.release();
    dref.release();
    if (rc != GPRT_OK) return rc;
    *best_out = best;
    *best_ms_out = best_ms;
    return GPRT_OK;
}

int finish_tune(const Cand& best, double best_ms, gprt_hip_config* cfg_out, int* leaf_size_out, int* build_flags_out,
                const char* json_path) {
    *cfg_out = best.cfg;
    *leaf_size_out = best.leaf;
    if (build_flags_out) *build_flags_out = best.sah ? GPRT_BUILD_SAH : 0;
    if (json_path) {
        FILE* f = fopen(json_path, "w");
        if (!f) return fail(GPRT_ERR_ARG, std::string("cannot write ") + json_path);
        const std::string js = schedule_json(best.cfg, best.leaf, best.sah, best_ms);
        fputs(js.c_str(), f);
        fclose(f);
    }
    return GPRT_OK;
}

}  // namespace

extern "C" int gprt_hip_autotune_ex(gprt_hip_ctx* h, const float* pts, int64_t n, const float* queries, int64_t nq,
                                    int k, int budget_trials, gprt_hip_config* cfg_out, int* leaf_size_out,
                                    int* build_flags_out, const char* json_path, int flags) {
    if (!h || !pts || !queries || !cfg_out || !leaf_size_out) return fail(GPRT_ERR_ARG, "null argument");
    GPRT_CTX_LOCK(h);
    if (n <= 0 || nq <= 0 || k < 1 || k > GPRT_HIP_K_MAX)
        return fail(GPRT_ERR_ARG, "need n > 0, nq > 0, 1 <= k <= " + std::to_string(GPRT_HIP_K_MAX));
    Ctx& c = h->c;
    GPRT_HIP_TRY(hipSetDevice(c.device));
    DevBuf dq;
    const float* d_q = queries;
    if (!(flags & GPRT_PTR_DEVICE)) {
        GPRT_TRY(dq.reserve(12ull * nq));
        GPRT_HIP_TRY(hipMemcpy(dq.ptr, queries, 12ull * nq, hipMemcpyHostToDevice));
        d_q = dq.as<float>();
    }
    Problem P;
    P.nout = (size_t)nq * k;
    P.build = [&](int leaf, int sah, gprt_hip_tree** out) {
        return gprt_hip_tree_build(h, pts, n, leaf, (flags & GPRT_PTR_DEVICE) | (sah ? GPRT_BUILD_SAH : 0), out);
    };
    P.run = [&](gprt_hip_tree* t, const gprt_hip_config& cfg, uint32_t* d_ids) {
        return gprt_hip_knn(h, t, d_q, nq, k, &cfg, d_ids, nullptr, GPRT_PTR_DEVICE);
    };
    Cand best{};
    double best_ms = 0;
    const int rc = autotune_core(h, P, budget_trials, build_flags_out != nullptr, &best, &best_ms);
    dq.release();
    if (rc != GPRT_OK) return rc;
    return finish_tune(best, best_ms, cfg_out, leaf_size_out, build_flags_out, json_path);
}

extern "C" int gprt_hip_autotune_batched(gprt_hip_ctx* h, const float* pts, const int64_t* cloud_offsets,
                                         int nclouds, const float* queries, const int64_t* query_offsets,
                                         const int* k_per_cloud, int budget_trials, gprt_hip_config* cfg_out,
                                         int* leaf_size_out, int* build_flags_out, const char* json_path,
                                         int flags) {
    if (!h || !pts || !cloud_offsets || !queries || !query_offsets || !k_per_cloud || !cfg_out || !leaf_size_out)
        return fail(GPRT_ERR_ARG, "null argument");
    if (nclouds < 1) return fail(GPRT_ERR_ARG, "nclouds must be >= 1");
    GPRT_CTX_LOCK(h);
    size_t nout = 0;
    for (int c = 0; c < nclouds; ++c) {
        const int64_t qn = query_offsets[c + 1] - query_offsets[c];
        if (qn < 0 || k_per_cloud[c] < 1 || k_per_cloud[c] > GPRT_HIP_K_MAX)
            return fail(GPRT_ERR_ARG, "bad query offsets or k of cloud " + std::to_string(c));
        nout += (size_t)qn * k_per_cloud[c];
    }
    const int64_t nq = query_offsets[nclouds] - query_offsets[0];
    if (query_offsets[0] != 0 || nq <= 0 || cloud_offsets[0] != 0 || cloud_offsets[nclouds] <= 0)
        return fail(GPRT_ERR_ARG, "need offsets from 0, some points and some queries");
    Ctx& c = h->c;
    GPRT_HIP_TRY(hipSetDevice(c.device));
    DevBuf dq;
    const float* d_q = queries;
    if (!(flags & GPRT_PTR_DEVICE)) {
        GPRT_TRY(dq.reserve(12ull * nq));
        GPRT_HIP_TRY(hipMemcpy(dq.ptr, queries, 12ull * nq, hipMemcpyHostToDevice));
        d_q = dq.as<float>();
    }
    Problem P;
    P.nout = nout;
    P.build = [&](int leaf, int sah, gprt_hip_tree** out) {
        return gprt_hip_tree_build_batched(h, pts, cloud_offsets, nclouds, leaf,
                                           (flags & GPRT_PTR_DEVICE) | (sah ? GPRT_BUILD_SAH : 0), out);
    };
    P.run = [&](gprt_hip_tree* t, const gprt_hip_config& cfg, uint32_t* d_ids) {
        return gprt_hip_knn_batched(h, t, d_q, query_offsets, k_per_cloud, &cfg, d_ids, nullptr, GPRT_PTR_DEVICE);
    };
    Cand best{};
    double best_ms = 0;
    const int rc = autotune_core(h, P, budget_trials, build_flags_out != nullptr, &best, &best_ms);
    dq.release();
    if (rc != GPRT_OK) return rc;
    return finish_tune(best, best_ms, cfg_out, leaf_size_out, build_flags_out, json_path);
}

extern "C" int gprt_hip_autotune(gprt_hip_ctx* h, const float* pts, int64_t n, const float* queries, int64_t nq,
                                 int k, int budget_trials, gprt_hip_config* cfg_out, int* leaf_size_out,
                                 const char* json_path, int flags) {
    return gprt_hip_autotune_ex(h, pts, n, queries, nq, k, budget_trials, cfg_out, leaf_size_out, nullptr, json_path,
                                flags);
}

// Schedule / wisdom JSON (gprt-ir/src/lib.rs:58-89) -> config.  Fields missing
// from the file take Schedule::default() (lib.rs:76-87): the reference's own
// *_wisdom.json files lack build_strategy and geom_type and therefore fail its
// serde reader; here they load.  max_hits_per_query, memory_strategy,
// build_strategy selects the tree (PreferFastTrace, Schedule::default()'s:
// the SAH build; PreferFastBuild: the LBVH) through *build_flags of the _ex
// form; geom_type has no counterpart (the points are spheres of radius 0).  An
// object "gprt_hip" (written by gprt_hip_autotune) sets the config fields it
// names and the leaf size.
extern "C" int gprt_hip_schedule_from_json_ex(const char* json, size_t n, gprt_hip_config* cfg, int* leaf_size,
                                              int* build_flags) {
    if (!json || !cfg) return fail(GPRT_ERR_ARG, "null argument");
    JParser jp(json, n);
    JVal v;
    if (!jp.value(v)) return fail(GPRT_ERR_ARG, "schedule: " + jp.err);
    if (v.kind != JVal::Obj) return fail(GPRT_ERR_ARG, "schedule: expected a JSON object");
    gprt_hip_config c;
    gprt_hip_config_default(&c);
    // Schedule::default()
    c.radius_heuristic = GPRT_RADIUS_PERCENTILE;
    c.radius_param = 0.10f;
    c.radius_mult = 3.0f;
    c.sort_queries = 0;
    int leaf = 8;
    auto num = [](const JVal* x, double& out) {
        if (!x || x->kind != JVal::Num) return false;
        out = x->num;
        return true;
    };
    double d = 0.0;
    auto finite_f32 = [&](const char* key) {
        if (!(std::fabs(d) <= 3.0e38)) {
            fail(GPRT_ERR_ARG, std::string("schedule: ") + key + " out of range");
            return false;
        }
        return true;
    };
    if (num(v.get("radius_increment_mult"), d)) {
        if (!finite_f32("radius_increment_mult")) return GPRT_ERR_ARG;
        c.radius_mult = (float)d;
    }
    if (const JVal* m = v.get("use_morton_lbv")) {
        if (m->kind != JVal::Bool) return fail(GPRT_ERR_ARG, "schedule: use_morton_lbv must be a boolean");
        c.sort_queries = m->b ? 1 : 0;
    }
    if (const JVal* h = v.get("radius_heuristic")) {
        if (h->kind == JVal::Str && h->str == "SampledMax") {
            c.radius_heuristic = GPRT_RADIUS_MAX;
        } else if (h->kind == JVal::Obj && num(h->get("SampledPercentile"), d)) {
            if (!finite_f32("SampledPercentile")) return GPRT_ERR_ARG;
            c.radius_heuristic = GPRT_RADIUS_PERCENTILE;
            c.radius_param = (float)d;
        } else if (h->kind == JVal::Obj && num(h->get("Fixed"), d)) {
            if (!finite_f32("Fixed")) return GPRT_ERR_ARG;
            c.radius_heuristic = GPRT_RADIUS_FIXED;
            c.radius_param = (float)d;
        } else {
            return fail(GPRT_ERR_ARG, "schedule: radius_heuristic must be \"SampledMax\", {\"SampledPercentile\": p} "
                                      "or {\"Fixed\": r}");
        }
    }
    int bflags = GPRT_BUILD_SAH;  // Schedule::default(): PreferFastTrace
    if (const JVal* bs = v.get("build_strategy")) {
        if (bs->kind == JVal::Str && bs->str == "PreferFastTrace") bflags = GPRT_BUILD_SAH;
        else if (bs->kind == JVal::Str && bs->str == "PreferFastBuild") bflags = 0;
        else return fail(GPRT_ERR_ARG, "schedule: build_strategy must be \"PreferFastTrace\" or \"PreferFastBuild\"");
    }
    if (const JVal* g = v.get("gprt_hip")) {
        if (g->kind != JVal::Obj) return fail(GPRT_ERR_ARG, "schedule: gprt_hip must be an object");
        struct {
            const char* key;
            int* ip;
            float* fp;
        } fields[] = {{"sampler", &c.sampler, nullptr},           {"radius_heuristic", &c.radius_heuristic, nullptr},
                      {"radius_param", nullptr, &c.radius_param}, {"radius_mult", nullptr, &c.radius_mult},
                      {"max_rounds", &c.max_rounds, nullptr},     {"traversal", &c.traversal, nullptr},
                      {"stack_depth", &c.stack_depth, nullptr},   {"guard_slots", &c.guard_slots, nullptr},
                      {"sort_queries", &c.sort_queries, nullptr}, {"use_graph", &c.use_graph, nullptr},
                      {"grid_blocks", &c.grid_blocks, nullptr},   {"lds_stack", &c.lds_stack, nullptr},
                      {"compact_rounds", &c.compact_rounds, nullptr}, {"leaf_size", &leaf, nullptr}};
        for (auto& f : fields)
            if (num(g->get(f.key), d)) {
                if (f.ip) {
                    if (!(d >= -2147483648.0 && d <= 2147483647.0) || d != (double)(int64_t)d)
                        return fail(GPRT_ERR_ARG, std::string("schedule: ") + f.key + " must be an integer");
                    *f.ip = (int)d;
                } else {
                    if (!(std::fabs(d) <= 3.0e38)) return fail(GPRT_ERR_ARG, std::string("schedule: ") + f.key + " out of range");
                    *f.fp = (float)d;
                }
            }
    }
    *cfg = c;
    if (leaf_size) *leaf_size = leaf;
    if (build_flags) *build_flags = bflags;
    return GPRT_OK;
}

extern "C" int gprt_hip_schedule_from_json(const char* json, size_t n, gprt_hip_config* cfg, int* leaf_size) {
    return gprt_hip_schedule_from_json_ex(json, n, cfg, leaf_size, nullptr);
}
