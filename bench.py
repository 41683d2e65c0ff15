"""Headline benchmark: exact k-NN, k = 5, on the LiDAR-like workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--leaf L] [--no-cpu] [--no-extra] [--all]

Workload: 1,310,720 float32 points (20 frames x 64 rings x 1,024 azimuths of a
ground-plane + boxes + 120 m cylinder scene, 2 cm range noise) and 400,000
queries (dataset points + 2 cm jitter), k = 5, generated with fixed seeds (the
sha256 of the generated arrays is part of the output).  One step = one exact
k-NN pass of all queries over a prebuilt LBVH, with the queries already
resident in HBM and the ids left resident in HBM.

Prints one JSON line.  `value` is queries/s of the batched FP32 + FP64-recheck
form with the default schedule (device-resident inputs, K steps enqueued back
to back between two synchronisations); the plain single-launch FP64 form is
timed the same way on the same tree.  The roofline entry is for the dominant
kernel (the first traversal round): algorithmic bytes counted by an
instrumented run, divided by the kernel's duration timed with HIP events on the
library stream.  The CPU baseline is the oracle's multi-threaded KD-tree on this
host's cores.  `details` adds, on the same data: the autotuner's pick, the
end-to-end time with the build and host transfers, k = 32 and the r = 0.5 m
range query (skipped with --no-extra); --all adds the road-network, batched
multi-cloud and Barnes-Hut workloads.
"""
import argparse
import hashlib
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rt-dsl_amd"))

import numpy as np  # noqa: E402

METRIC = "exact kNN queries/s, k=5, 1.3M pts: batched FP32+FP64-recheck vs plain GPU"
HBM_PEAK_GBS = 8000.0
K = 5


def _sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def _time_steps(ctx, fn, steps, warmup):
    for _ in range(warmup):
        fn(False)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn(False)
    ctx.sync()
    return (time.perf_counter() - t0) * 1e3 / steps


def _cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return platform.processor() or "unknown"


def _knn_pair(ctx, tree, dq, k, cfg, steps, warmup):
    """(batched ms/step, plain ms/step, batched ids, plain ids) on device arrays."""
    import gprt_hip
    nq = len(dq)
    ib = gprt_hip.DeviceArray(ctx, (nq, k), np.uint32)
    ip = gprt_hip.DeviceArray(ctx, (nq, k), np.uint32)
    ms_b = _time_steps(ctx, lambda sync: tree.k_nn(dq, k, cfg, ids_out=ib, sync=sync), steps, warmup)
    ms_p = _time_steps(ctx, lambda sync: tree.k_nn_plain(dq, k, ids_out=ip, sync=sync), steps, warmup)
    tree.k_nn(dq, k, cfg, ids_out=ib)
    st = ctx.stats()
    out = ib.to_host(), ip.to_host()
    ib.free()
    ip.free()
    return ms_b, ms_p, out[0], out[1], st


def _extra_lidar(ctx, tree, pts, q, dq, cfg, args, oracle):
    """k = 32 and the r = 0.5 m range query on the headline data."""
    ex = {}
    steps = max(5, args.steps // 5)
    ms_b, ms_p, ib, ip, st = _knn_pair(ctx, tree, dq, 32, cfg, steps, 2)
    ref, _, _, _ = oracle.kdtree_knn(pts, q, 32)
    ex["knn_k32"] = {"batched_qps": len(q) / (ms_b * 1e-3), "plain_qps": len(q) / (ms_p * 1e-3),
                     "speedup_vs_plain": ms_p / ms_b, "ms_batched": ms_b, "ms_plain": ms_p,
                     "fallback": st["fallback_queries"], "rounds": st["rounds_run"],
                     "batched_eq_oracle": bool(np.array_equal(ib, ref)),
                     "plain_eq_oracle": bool(np.array_equal(ip, ref))}
    # range query r = 0.5 m: device time of the fill call (Morton order, FP32
    # traversal with FP64 confirmation, CSR write, per-query (d64, id) sort)
    t0 = time.perf_counter()
    off, ids = tree.range_query(q, 0.5)
    wall = time.perf_counter() - t0
    ms = ctx.stats()["ms_query"]
    lens = np.diff(off)
    ok_sorted = True
    # spot-check 2,000 queries against the oracle's brute force (every hit within r, (d64, id) order, complete)
    rng = np.random.default_rng(11)
    sel = np.sort(rng.choice(len(q), 2000, replace=False))
    roff, rids = oracle.range_brute(pts, q[sel], 0.5)
    got = np.concatenate([ids[off[i]:off[i + 1]] for i in sel])
    ok_sorted = bool(np.array_equal(np.diff(roff), lens[sel]) and np.array_equal(got, rids))
    ex["range_r0.5"] = {"qps": len(q) / (ms * 1e-3), "ms_device": ms, "wall_s_incl_transfers": wall,
                        "results": int(off[-1]), "mean_per_query": float(lens.mean()), "max_per_query": int(lens.max()),
                        "results_per_s": float(off[-1]) / (ms * 1e-3),
                        "sample_eq_oracle": ok_sorted, "sample": "2,000 random queries vs FP64 brute force"}
    return ex


def _extra_all(ctx, args, oracle):
    """Road network (plain vs CPU), batched multi-cloud, Barnes-Hut."""
    import gprt_hip
    from gprt_hip import datasets as D
    ex = {}
    steps = max(5, args.steps // 5)
    # road network: 434K points x 400K queries
    pts, q = D.road_network()
    dp, dq = gprt_hip.DeviceArray.from_host(ctx, pts), gprt_hip.DeviceArray.from_host(ctx, q)
    tree = ctx.build(dp, leaf_size=8)
    ms_b, ms_p, ib, ip, st = _knn_pair(ctx, tree, dq, K, gprt_hip.default_config(), steps, 2)
    ref, _, b_ms, q_ms = oracle.kdtree_knn(pts, q, K)
    ex["road_network"] = {"sha256_16": _sha(pts, q), "batched_qps": len(q) / (ms_b * 1e-3),
                          "plain_qps": len(q) / (ms_p * 1e-3), "cpu_qps": len(q) / (q_ms * 1e-3),
                          "batched_eq_oracle": bool(np.array_equal(ib, ref)),
                          "plain_eq_oracle": bool(np.array_equal(ip, ref))}
    tree.close()
    dp.free()
    dq.free()
    # batched multi-cloud: 256 clouds, k in {1, 5, 16, 64}
    pts, poff, q, qoff, kc = D.clustered_batch()
    tree = ctx.build_batched(pts, poff, leaf_size=8)
    cfg = gprt_hip.default_config()
    ids, out_off = tree.k_nn_batched(q, qoff, kc, cfg)
    times = []
    for _ in range(steps):
        tree.k_nn_batched(q, qoff, kc, cfg)
        times.append(ctx.stats()["ms_query"])
    ms = float(np.median(times))
    # parity on 16 clouds spread over the size range (per-cloud KD-tree oracle)
    ok = True
    for c in np.linspace(0, len(kc) - 1, 16).astype(int):
        k = int(kc[c])
        r, _, _, _ = oracle.kdtree_knn(pts[poff[c]:poff[c + 1]], q[qoff[c]:qoff[c + 1]], k)
        got = ids[out_off[c]:out_off[c + 1]].reshape(-1, k)
        ok &= bool(np.array_equal(got, r))
    ex["batched_clouds"] = {"clouds": len(kc), "points": len(pts), "queries": len(q), "ms_device": ms,
                            "qps": len(q) / (ms * 1e-3), "sample_eq_oracle": ok,
                            "sample": "16 clouds across the size range vs the KD-tree oracle"}
    tree.close()
    # Barnes-Hut: 1M-body Plummer sphere, theta in {0.3, 0.5, 0.7}, error on 10,000 sampled bodies
    pm = D.plummer()
    dpm = gprt_hip.DeviceArray.from_host(ctx, pm)
    sample = np.sort(np.random.default_rng(5).choice(len(pm), 10_000, replace=False))
    ref = oracle.bh_direct(pm, sample, G=1.0, eps=1e-3)
    bh = {}
    for theta in (0.3, 0.5, 0.7):
        acc = ctx.barnes_hut(dpm, theta=theta, eps=1e-3, G=1.0)
        st = ctx.stats()
        a = acc.to_host()[sample].astype(np.float64)
        err = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
        bh[str(theta)] = {"ms_force": st["ms_query"], "ms_total": st["ms_total"], "bodies_per_s": len(pm) / (st["ms_query"] * 1e-3),
                          "rel_err_median": float(np.median(err)), "rel_err_p99": float(np.percentile(err, 99)),
                          "rel_err_max": float(err.max())}
        acc.free()
    ex["barnes_hut_plummer_1m"] = bh
    dpm.free()
    return ex


def _only(args):
    """One secondary workload alone, for profiling: the r = 0.5 m range query on
    the LiDAR data, or Barnes-Hut on the 1M-body Plummer sphere.  Prints one
    JSON line for that workload (not the headline metric)."""
    import gprt_hip
    from gprt_hip import datasets as D
    ctx = gprt_hip.Context(0)
    steps = max(2, args.steps)
    if args.only == "range":
        pts, q = D.lidar_scan()
        tree = ctx.build(pts, leaf_size=args.leaf)
        ms = []
        for _ in range(steps + 1):
            off, ids = tree.range_query(q, 0.5)
            ms.append(ctx.stats()["ms_query"])
        line = {"workload": "lidar_scan range r=0.5 m, 1,310,720 pts x 400,000 queries", "steps": steps,
                "ms_device_fill_call": float(np.median(ms[1:])), "results": int(off[-1])}
    else:
        pm = D.plummer()
        dpm = gprt_hip.DeviceArray.from_host(ctx, pm)
        line = {"workload": "barnes-hut plummer 1,000,000 bodies, eps=1e-3, G=1", "steps": steps}
        for theta in (0.3, 0.5, 0.7):
            ms = []
            for _ in range(steps + 1):
                ctx.barnes_hut(dpm, theta=theta, eps=1e-3, G=1.0).free()
                ms.append(ctx.stats()["ms_query"])
            line[f"ms_force_theta_{theta}"] = float(np.median(ms[1:]))
    print(json.dumps(line))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--leaf", type=int, default=8, help="leaf size of the LBVH")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip autotune, end-to-end, k=32 and range")
    ap.add_argument("--all", action="store_true", help="also the road, batched-cloud and Barnes-Hut workloads")
    ap.add_argument("--only", choices=["range", "barnes-hut"],
                    help="run one secondary workload alone (for profiling) and print its JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the k-NN path would shard queries over a replicated tree)", file=sys.stderr)
        return 2
    if args.only:
        return _only(args)

    import gprt_hip
    from gprt_hip import datasets as D
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    t0 = time.perf_counter()
    pts, q = D.lidar_scan()
    gen_s = time.perf_counter() - t0
    nq = len(q)
    ctx = gprt_hip.Context(0)
    dp = gprt_hip.DeviceArray.from_host(ctx, pts)
    dq = gprt_hip.DeviceArray.from_host(ctx, q)
    ids_b = gprt_hip.DeviceArray(ctx, (nq, K), np.uint32)
    ids_p = gprt_hip.DeviceArray(ctx, (nq, K), np.uint32)

    leaf = args.leaf
    cfg = gprt_hip.default_config()
    cfg.use_graph = 1
    tree = ctx.build(dp, leaf_size=leaf)
    build_ms = ctx.stats()["ms_build"]

    def batched(sync):
        tree.k_nn(dq, K, cfg, ids_out=ids_b, sync=sync)

    def plain(sync):
        tree.k_nn_plain(dq, K, ids_out=ids_p, sync=sync)

    ms_b = _time_steps(ctx, batched, args.steps, args.warmup)
    batched(True)
    st_b = ctx.stats()
    ms_p = _time_steps(ctx, plain, args.steps, args.warmup)
    plain(True)
    ms_p_dev = ctx.stats()["ms_query"]

    # roofline of the dominant kernel (first traversal round): counts from an
    # instrumented run, duration from HIP events around the uninstrumented kernel
    ccfg = gprt_hip.default_config(**cfg.as_dict())
    ccfg.use_graph, ccfg.collect_counters = 0, 1
    tree.k_nn(dq, K, ccfg, ids_out=ids_b)
    cnt = ctx.stats()
    tcfg = gprt_hip.default_config(**cfg.as_dict())
    tcfg.use_graph, tcfg.time_kernels = 0, 1
    kms = []
    for _ in range(max(5, args.steps // 5)):
        tree.k_nn(dq, K, tcfg, ids_out=ids_b)
        kms.append(ctx.stats()["ms_kernel"])
    k_ms = float(np.mean(kms))
    r0 = cnt["active_per_round"][0]
    # per query: 16 B query + 32 B per AABB test (two per node visit) + 16 B
    # per point test; + 4k B of ids written + 16 B per FP64 re-rank read
    alg_bytes = 16 * r0 + 32 * 2 * cnt["node_visits"] + 16 * cnt["point_tests"] + 4 * K * nq + 16 * cnt["rerank_reads"]
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9

    # a device-to-device copy of known size: the reference point for scaling
    # rocprofv3 FETCH_SIZE / WRITE_SIZE counters to bytes
    copy_bytes = 64 << 20
    cbuf_a = gprt_hip.DeviceArray(ctx, (copy_bytes // 4,), np.uint32)
    cbuf_b = gprt_hip.DeviceArray(ctx, (copy_bytes // 4,), np.uint32)
    gprt_hip.lib().gprt_hip_copy(ctx.h, cbuf_b.ptr, cbuf_a.ptr, copy_bytes, 2)
    ctx.sync()
    cbuf_a.free()
    cbuf_b.free()
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "knn_round0_lidar_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    # parity of both forms against the oracle (CPU KD-tree, pinned to brute force)
    batched(True)
    got_b = ids_b.to_host()
    got_p = ids_p.to_host()
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    runs = []
    ref = None
    for i in range(6):  # one warm-up, then the median of five
        ref, _, b_ms, q_ms = oracle.kdtree_knn(pts, q, K, nthreads=nthreads)
        if i:
            runs.append((b_ms, q_ms))
    q_med = float(np.median([r[1] for r in runs]))
    b_med = float(np.median([r[0] for r in runs]))
    parity = {"batched_eq_plain": bool(np.array_equal(got_b, got_p)),
              "batched_eq_oracle": bool(np.array_equal(got_b, ref)),
              "plain_eq_oracle": bool(np.array_equal(got_p, ref))}
    cpu = None if args.no_cpu else {
        "value": nq / (q_med * 1e-3), "unit": "queries/s", "cores": nthreads, "kind": "port",
        "sample": f"full workload: {nq} queries over {len(pts)} points, k={K}; oracle KD-tree "
                  f"(median split, leaf 16, FP64, OpenMP over queries), one warm-up then the median of 5 query "
                  f"phases ({q_med:.1f} ms); tree build {b_med:.1f} ms excluded",
        "nproc": os.cpu_count(), "cpu_model": _cpu_model()}

    details = {"ms_device_batched": st_b["ms_query"], "rounds": st_b["rounds_run"],
               "active_per_round": st_b["active_per_round"][:8], "fallback": st_b["fallback_queries"],
               "build_ms": build_ms, "datagen_s": gen_s, "data_sha256_16": _sha(pts, q),
               "seeds": {"lidar_scan": 3}}
    if not args.no_extra:
        # the autotuner's pick on this data (every candidate must reproduce the
        # default schedule's ids); the headline keeps the default schedule
        t0 = time.perf_counter()
        acfg, aleaf = ctx.autotune(dp, dq, K, budget_trials=14)
        details["autotune"] = {"seconds": time.perf_counter() - t0, "leaf_size": aleaf, "schedule": acfg.as_dict()}
        # end to end: host points -> build -> host queries -> ids on the host
        e2e = []
        for _ in range(3):
            t0 = time.perf_counter()
            t2 = ctx.build(pts, leaf_size=leaf)
            t2.k_nn(q, K, cfg)
            e2e.append(time.perf_counter() - t0)
            t2.close()
        details["end_to_end_ms"] = {"median": 1e3 * float(np.median(e2e)),
                                    "what": "H2D points + LBVH build + H2D queries + batched k-NN + D2H ids, "
                                            "wall clock"}
        details.update(_extra_lidar(ctx, tree, pts, q, dq, cfg, args, oracle))
    if args.all:
        details.update(_extra_all(ctx, args, oracle))

    out = {
        "metric": METRIC,
        "value": nq / (ms_b * 1e-3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_b,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "lidar_scan: 1,310,720 pts (20x64x1024) x 400,000 queries, k=5",
                   "leaf_size": leaf, "schedule": cfg.as_dict(), "form": "batched FP32 + FP64 re-rank"},
        "plain": {"value": nq / (ms_p * 1e-3), "ms_per_step": ms_p, "ms_device": ms_p_dev,
                  "form": "one launch, one query per lane in input order, FP64, r=inf"},
        "speedup_vs_plain": ms_p / ms_b,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "knn round 0 traversal", "kernel_ms": k_ms, "alg_bytes": int(alg_bytes),
                     "node_visits": cnt["node_visits"], "point_tests": cnt["point_tests"],
                     "rerank_reads": cnt["rerank_reads"],
                     "node_visits_per_s": cnt["node_visits"] / (k_ms * 1e-3)},
        "cpu_baseline": cpu,
        "parity": parity,
        "details": details,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
