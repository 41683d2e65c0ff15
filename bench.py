"""Headline benchmark: exact k-NN, k = 5, on the LiDAR-like workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--leaf L] [--no-cpu] [--no-extra] [--all]
                    [--traffic auto|off] [--only range|barnes-hut|build] [--sah]

Workload: 1,310,720 float32 points (20 frames x 64 rings x 1,024 azimuths of a
ground-plane + boxes + 120 m cylinder scene, 2 cm range noise) and 400,000
queries (dataset points + 2 cm jitter), k = 5, generated with fixed seeds (the
sha256 of the generated arrays is part of the output).  One step = one exact
k-NN pass of all queries over a prebuilt tree, with the queries already
resident in HBM and the ids left resident in HBM.

Prints one JSON line.  `value` is queries/s of the batched FP32 + FP64-recheck
form (device-resident inputs, K steps enqueued back to back between two
synchronisations), with the default schedule (LBVH, leaf 8) or the autotuner's
pick (schedule, leaf size and tree: LBVH or SAH build) when that is faster at
identical ids; `config` names the tree and its build time.  The plain
single-launch FP64 form is timed on the default tree.  The roofline entry is
for the dominant kernel (the first traversal round) of the configuration
`value` used: algorithmic bytes counted by an instrumented run, divided by the
kernel's duration timed with HIP events on the library stream; `traffic` is
the kernel's HBM bytes per launch from two rocprofv3 --pmc passes (FETCH_SIZE,
WRITE_SIZE) over the same configuration in child processes of this run, scaled
by what the same counters read for a 64 MiB streaming copy kernel in those
passes (null when rocprofv3 is absent or --traffic off).  The CPU baseline is the oracle's multi-threaded KD-tree on this
host's cores.  `details` adds, on the same data: the end-to-end time with the
build and host transfers, k = 32, the r = 0.5 m range query, the LBVH
rebuild and the SAH build, with rooflines (skipped with --no-extra); --all adds the
road-network, batched multi-cloud and Barnes-Hut workloads.  --only runs one
secondary workload alone and prints its own line (for profiling).
"""
import argparse
import csv
import glob
import hashlib
import json
import os
import platform
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rt-dsl_amd"))

import numpy as np  # noqa: E402

METRIC = "exact kNN queries/s, k=5, 1.3M pts: batched FP32+FP64-recheck vs plain GPU"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
K = 5
ROUND0 = "knn_lane_kernel<8, 3, 8, false, true>"
COPY_BYTES = 64 << 20


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


def _roof(alg_bytes, ms, **extra):
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    return dict({"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": achieved / HBM_PEAK_GBS, "alg_bytes": int(alg_bytes), "kernel_ms": ms}, **extra)


# ---------------------------------------------------------------- HBM traffic
def _pmc_pass(counter, tmp, head):
    """One rocprofv3 --pmc pass over `bench.py --only knn-round0` in a child
    process running the head configuration `head` (leaf size, tree, schedule);
    returns {kernel name: mean counter value (KiB) per dispatch}."""
    d = os.path.join(tmp, counter.lower())
    hargs = ["--leaf", str(head["leaf"]), "--schedule", json.dumps(head["schedule"])] + (["--sah"] if head["sah"] else [])
    cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--",
           sys.executable, os.path.abspath(__file__), "--only", "knn-round0", "--traffic", "off"] + hargs
    env = dict(os.environ, TMPDIR=tmp)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=180, env=env, cwd=tmp)
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"{counter}: no counter file (rc {r.returncode}): {r.stderr[-400:]}")
    acc = {}
    with open(files[0]) as f:
        for row in csv.DictReader(f):
            if row.get("Counter_Name") != counter:
                continue
            acc.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    shown = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", "<tmp>", "-o", "run", "--", "python3",
             "bench.py", "--only", "knn-round0", "--traffic", "off", "--leaf", str(head["leaf"])] + \
        (["--sah"] if head["sah"] else []) + ["--schedule", "'<head schedule>'"]
    return {k: sum(v) / len(v) for k, v in acc.items()}, " ".join(shown)


def _traffic(head):
    """HBM bytes per launch of the first traversal round of the head
    configuration from two counter passes, each scaled by the counter's reading
    for a plain 64 MiB copy."""
    if shutil.which("rocprofv3") is None:
        return None, "rocprofv3 not found"
    tmp = tempfile.mkdtemp(prefix="gprt_pmc_")
    try:
        fetch, c1 = _pmc_pass("FETCH_SIZE", tmp, head)
        write, c2 = _pmc_pass("WRITE_SIZE", tmp, head)
        kern = [k for k in fetch if ROUND0 in k]
        copy = [k for k in fetch if "stream_copy_kernel" in k]
        if not kern or not copy:
            return None, f"kernels not found in the counter files: {sorted(fetch)[:6]}"
        kib = COPY_BYTES / 1024.0
        fs = kib / fetch[copy[0]]     # the copy reads COPY_BYTES
        ws = kib / write[copy[0]]     # and writes COPY_BYTES
        t = {"hbm_bytes_per_launch": 1024.0 * (fetch[kern[0]] * fs + write[kern[0]] * ws),
             "fetch_kib": fetch[kern[0]], "write_kib": write[kern[0]], "fetch_scale": fs, "write_scale": ws,
             "write_bytes_per_launch": 1024.0 * write[kern[0]] * ws, "commands": [c1, c2]}
        return t, None
    except Exception as e:  # the line still prints; the reason is recorded
        return None, f"{type(e).__name__}: {e}"
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


# ---------------------------------------------------------------- k-NN helpers
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


def _range_line(ctx, tree, pts, q, oracle, reps=3):
    """r = 0.5 m over the LiDAR scan: both calls timed (count, then fill and
    per-query sort), the fill walk alone for the roofline, counts from an
    instrumented run, 2,000 sampled queries against FP64 brute force."""
    import gprt_hip
    nq = len(q)
    tot, cnt_ms, fill_ms, walk_ms, cwalk_ms = [], [], [], [], []
    off = ids = None
    for _ in range(reps + 1):
        st = []
        off, ids = tree.range_query(q, 0.5, flags=gprt_hip.TIME_KERNELS, stats=st)
        cnt_ms.append(st[0]["ms_query"])
        fill_ms.append(st[1]["ms_query"])
        cwalk_ms.append(st[0]["ms_kernel"])
        walk_ms.append(st[1]["ms_kernel"])
        tot.append(st[0]["ms_query"] + st[1]["ms_query"])
    med = lambda v: float(np.median(v[1:]))  # noqa: E731
    st = []
    tree.range_query(q, 0.5, flags=gprt_hip.COLLECT_COUNTERS, stats=st)
    V, P, S = st[1]["node_visits"], st[1]["point_tests"], st[1]["rerank_reads"]
    hits = int(off[-1])
    # fill walk: 16 B query + 64 B per node visit + 16 B per point test + 4 B
    # per output slot written (a hit's position, or two per subtree taken
    # whole) + 16 B of offsets per query
    alg = 16 * nq + 64 * V + 16 * P + 4 * S + 16 * nq
    rng = np.random.default_rng(11)
    sel = np.sort(rng.choice(nq, 2000, replace=False))
    roff, rids = oracle.range_brute(pts, q[sel], 0.5)
    got = np.concatenate([ids[off[i]:off[i + 1]] for i in sel])
    ok = bool(np.array_equal(np.diff(off)[sel], np.diff(roff)) and np.array_equal(got, rids))
    return {"workload": "lidar_scan range r=0.5 m, 1,310,720 pts x 400,000 queries",
            "ms_count_plus_fill": med(tot), "ms_count_call": med(cnt_ms), "ms_fill_call": med(fill_ms),
            "qps": nq / (med(tot) * 1e-3), "results": hits, "results_per_s": hits / (med(tot) * 1e-3),
            "mean_per_query": hits / nq, "max_per_query": int(np.diff(off).max()),
            "roofline": _roof(alg, med(walk_ms), kernel="range fill walk (range_kernel<true>)", node_visits=V,
                              point_tests=P, slots=S, count_walk_ms=med(cwalk_ms)),
            "sample_eq_oracle": ok, "sample": "2,000 random queries vs FP64 brute force"}


def _build_line(ctx, dp, n, leaf, reps=7):
    """LBVH rebuild from device-resident points (the context keeps the last
    destroyed tree's buffers).  Bytes per phase: bounds 12N read, keys 12N
    read + 12N written, 8 radix passes over (8 B key, 4 B value) 24N each,
    gather 16N read + 16N + 16N written (sorted and input-order float4),
    Karras ~8 log2(N) N read + 40N written, refit 16N + 48N, 4-wide view 64N
    read + 128N written."""
    ms = []
    for _ in range(reps + 1):
        t = ctx.build(dp, leaf_size=leaf)
        ms.append(ctx.stats()["ms_build"])
        t.close()
    med = float(np.median(ms[1:]))
    alg = n * (12 + 24 + 8 * 24 + 48 + 8 * np.log2(n) + 40 + 64 + 192)
    return {"workload": f"LBVH rebuild, {n} device-resident points, leaf {leaf}", "ms": med,
            "roofline": _roof(alg, med, kernel="LBVH build (all kernels)")}


def _sah_build_line(ctx, dp, n, leaf, reps=5):
    """SAH tree build (GPRT_BUILD_SAH) from device-resident points: the LBVH's
    Morton sort, then the SAH levels, the refit and the 4-wide view.  Build
    time only (its passes are many small latency-bound launches)."""
    ms = []
    for _ in range(reps + 1):
        t = ctx.build(dp, leaf_size=leaf, sah=True)
        ms.append(ctx.stats()["ms_build"])
        t.close()
    return {"workload": f"SAH tree build, {n} device-resident points, leaf {leaf}", "ms": float(np.median(ms[1:]))}


def _bh_line(ctx, pm, oracle, thetas=(0.3, 0.5, 0.7), check=True):
    """1M Plummer bodies: force kernel time per theta (HIP events), counts from
    an instrumented run.  The kernel walks the tree once per wavefront of 64
    Morton-adjacent bodies, so its bytes are per wavefront: 96 per node it
    loads (64 B node + 32 B moments) + 20 per bucket body it loads (float4 +
    mass), plus per body 16 read and 12 written.  The per-body count of the
    same walks (96 per node opened + 20 per body interaction, as if every body
    walked alone) is reported beside it.  FP32 work = 30 flops per interaction
    (body or far node: 3 sub, 3 fma, sqrt, div, 4 mul, 3 mul, 3 compensated
    adds of 4) + 32 per node opened (two opening tests)."""
    import gprt_hip
    n = len(pm)
    dpm = gprt_hip.DeviceArray.from_host(ctx, pm)
    out = {"workload": "barnes-hut plummer 1,000,000 bodies, eps=1e-3, G=1, leaf 8"}
    ref = sample = None
    if check:
        sample = np.sort(np.random.default_rng(5).choice(n, 10_000, replace=False))
        ref = oracle.bh_direct(pm, sample, G=1.0, eps=1e-3)
    for theta in thetas:
        ms = []
        for _ in range(4):
            ctx.barnes_hut(dpm, theta=theta, eps=1e-3, G=1.0).free()
            ms.append(ctx.stats()["ms_query"])
        acc = ctx.barnes_hut(dpm, theta=theta, eps=1e-3, G=1.0, flags=gprt_hip.COLLECT_COUNTERS)
        st = ctx.stats()
        V, L, F = st["node_visits"], st["point_tests"], st["rerank_reads"]
        Vw, Lw = st["wave_steps"], st["lane_steps"]
        t = float(np.median(ms[1:]))
        alg = 16 * n + 96 * Vw + 20 * Lw + 12 * n
        per_body = 16 * n + 96 * V + 20 * L + 12 * n
        flops = 30 * (L + F) + 32 * V
        line = {"ms_force": t, "bodies_per_s": n / (t * 1e-3),
                "roofline": _roof(alg, t, kernel="bh_wave_kernel (one walk per wavefront of 64 bodies)",
                                  wave_node_loads=Vw, wave_body_loads=Lw, nodes_opened=V, body_interactions=L,
                                  far_interactions=F, per_body_walk_bytes=per_body,
                                  fp32_tflops=flops / (t * 1e-3) / 1e12,
                                  fp32_frac=flops / (t * 1e-3) / 1e12 / FP32_PEAK_TFLOPS)}
        if check:
            a = acc.to_host()[sample].astype(np.float64)
            err = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
            line.update(rel_err_median=float(np.median(err)), rel_err_p99=float(np.percentile(err, 99)),
                        rel_err_max=float(err.max()))
        acc.free()
        out[str(theta)] = line
    dpm.free()
    return out


def _extra_all(ctx, args, oracle):
    """Road network (plain vs CPU), batched multi-cloud, Barnes-Hut."""
    import gprt_hip
    from gprt_hip import datasets as D
    ex = {}
    steps = max(5, args.steps // 5)
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
    pts, poff, q, qoff, kc = D.clustered_batch()

    def clouds_run(tree, cfg):
        ids, out_off = tree.k_nn_batched(q, qoff, kc, cfg)
        times = []
        for _ in range(steps):
            tree.k_nn_batched(q, qoff, kc, cfg)
            times.append(ctx.stats()["ms_query"])
        return ids, out_off, float(np.median(times))

    tree = ctx.build_batched(pts, poff, leaf_size=8)
    ids, out_off, ms_default = clouds_run(tree, gprt_hip.default_config())
    tree.close()
    # the batched autotuner's pick (schedule, leaf size and tree), adopted when
    # faster at identical ids
    t0 = time.perf_counter()
    tcfg, tleaf, tsah = ctx.autotune_batched(pts, poff, q, qoff, kc, budget_trials=24, with_tree=True)
    tune_s = time.perf_counter() - t0
    tree = ctx.build_batched(pts, poff, leaf_size=tleaf, sah=tsah)
    tids, _, ms_tuned = clouds_run(tree, tcfg)
    adopt = bool(np.array_equal(tids, ids)) and ms_tuned < 0.99 * ms_default
    if adopt:
        ids = tids
    ms = ms_tuned if adopt else ms_default
    ok = True
    for c in range(len(kc)):
        k = int(kc[c])
        r, _, _, _ = oracle.kdtree_knn(pts[poff[c]:poff[c + 1]], q[qoff[c]:qoff[c + 1]], k)
        ok &= bool(np.array_equal(ids[out_off[c]:out_off[c + 1]].reshape(-1, k), r))
    ex["batched_clouds"] = {"clouds": len(kc), "points": len(pts), "queries": len(q), "ms_device": ms,
                            "qps": len(q) / (ms * 1e-3), "all_clouds_eq_oracle": ok,
                            "ms_default_schedule": ms_default,
                            "autotune": {"seconds": tune_s, "leaf_size": tleaf, "tree": "sah" if tsah else "lbvh",
                                         "schedule": tcfg.as_dict(), "ms": ms_tuned, "adopted": adopt}}
    tree.close()
    ex["barnes_hut_plummer_1m"] = _bh_line(ctx, D.plummer(), oracle)
    return ex


def _only(args):
    """One workload alone (for profiling): the r = 0.5 m range query on the
    LiDAR data, Barnes-Hut on the 1M-body Plummer sphere, the LBVH rebuild, or
    (knn-round0, used by the counter passes) the first traversal round of the
    headline alone plus a 64 MiB streaming copy (gprt_hip_debug_stream_copy)."""
    import gprt_hip
    from gprt_hip import datasets as D
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    ctx = gprt_hip.Context(0)
    if args.only == "range":
        pts, q = D.lidar_scan()
        tree = ctx.build(pts, leaf_size=args.leaf, sah=args.sah)
        line = _range_line(ctx, tree, pts, q, oracle, reps=max(2, args.steps))
    elif args.only == "build":
        pts, _ = D.lidar_scan()
        dp = gprt_hip.DeviceArray.from_host(ctx, pts)
        line = _sah_build_line(ctx, dp, len(pts), args.leaf, reps=max(3, args.steps)) if args.sah else \
            _build_line(ctx, dp, len(pts), args.leaf, reps=max(3, args.steps))
    elif args.only == "knn-round0":
        pts, q = D.lidar_scan()
        dp, dq = gprt_hip.DeviceArray.from_host(ctx, pts), gprt_hip.DeviceArray.from_host(ctx, q)
        ids = gprt_hip.DeviceArray(ctx, (len(q), K), np.uint32)
        tree = ctx.build(dp, leaf_size=args.leaf, sah=args.sah)
        cfg = gprt_hip.default_config(**(json.loads(args.schedule) if args.schedule else {}))
        cfg.use_graph = 0
        for _ in range(max(2, args.steps)):
            tree.k_nn(dq, K, cfg, ids_out=ids)
        a = gprt_hip.DeviceArray(ctx, (COPY_BYTES // 4,), np.uint32)
        b = gprt_hip.DeviceArray(ctx, (COPY_BYTES // 4,), np.uint32)
        for _ in range(3):
            gprt_hip.lib().gprt_hip_debug_stream_copy(ctx.h, b.ptr, a.ptr, COPY_BYTES)
        line = {"workload": "knn round 0 (lidar_scan, k=5) and a 64 MiB device copy, for counter passes"}
    else:
        line = _bh_line(ctx, D.plummer(), oracle, check=not args.no_cpu)
    print(json.dumps(line))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--leaf", type=int, default=8, help="leaf size of the LBVH")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip autotune, end-to-end, k=32, range and build")
    ap.add_argument("--all", action="store_true", help="also the road, batched-cloud and Barnes-Hut workloads")
    ap.add_argument("--traffic", choices=["auto", "off"], default="auto",
                    help="measure the round kernel's HBM bytes with two rocprofv3 --pmc child passes")
    ap.add_argument("--only", choices=["range", "barnes-hut", "build", "knn-round0"],
                    help="run one workload alone (for profiling) and print its JSON line")
    ap.add_argument("--sah", action="store_true",
                    help="build the SAH tree (GPRT_BUILD_SAH) as the default tree (with --only: its tree)")
    ap.add_argument("--schedule", default=None, help="with --only knn-round0: the schedule as a JSON object")
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
    tree = ctx.build(dp, leaf_size=leaf, sah=args.sah)
    build_ms = ctx.stats()["ms_build"]

    def batched_on(t, c):
        return lambda sync: t.k_nn(dq, K, c, ids_out=ids_b, sync=sync)

    ms_b = _time_steps(ctx, batched_on(tree, cfg), args.steps, args.warmup)
    tree.k_nn(dq, K, cfg, ids_out=ids_b)
    st_b = ctx.stats()
    got_default = ids_b.to_host()
    ms_p = _time_steps(ctx, lambda sync: tree.k_nn_plain(dq, K, ids_out=ids_p, sync=sync), args.steps,
                       args.warmup)
    tree.k_nn_plain(dq, K, ids_out=ids_p)
    ms_p_dev = ctx.stats()["ms_query"]

    # the autotuner's pick (schedule, leaf size and tree: LBVH or SAH), timed
    # the same way; adopted when faster at equal ids
    auto = None
    head_tree, head_cfg, head_leaf, head_sah, head_ms, head_build_ms = tree, cfg, leaf, args.sah, ms_b, build_ms
    if not args.no_extra:
        t0 = time.perf_counter()
        acfg, aleaf, asah = ctx.autotune(dp, dq, K, budget_trials=16, with_tree=True)
        tune_s = time.perf_counter() - t0
        if aleaf == leaf and asah == args.sah:
            atree, abuild_ms = tree, build_ms
        else:
            atree = ctx.build(dp, leaf_size=aleaf, sah=asah)
            abuild_ms = ctx.stats()["ms_build"]
        ms_a = _time_steps(ctx, batched_on(atree, acfg), args.steps, args.warmup)
        atree.k_nn(dq, K, acfg, ids_out=ids_b)
        same = bool(np.array_equal(ids_b.to_host(), got_default))
        adopt = same and ms_a < ms_b * 0.99
        auto = {"seconds": tune_s, "leaf_size": aleaf, "tree": "sah" if asah else "lbvh", "build_ms": abuild_ms,
                "schedule": acfg.as_dict(), "ms_per_step": ms_a,
                "default_ms_per_step": ms_b, "ids_equal_default": same, "adopted": adopt}
        if adopt:
            head_tree, head_cfg, head_leaf, head_sah, head_ms, head_build_ms = atree, acfg, aleaf, asah, ms_a, abuild_ms
        elif atree is not tree:
            atree.close()

    # HBM traffic of the head configuration's round kernel: two counter passes
    # in child processes (this process idles meanwhile)
    ctx.sync()
    traffic, traffic_note = (None, "--traffic off") if args.traffic == "off" else _traffic(
        {"leaf": head_leaf, "sah": head_sah, "schedule": head_cfg.as_dict()})

    # roofline of the dominant kernel (first traversal round): counts from an
    # instrumented run, duration from HIP events around the uninstrumented kernel
    ccfg = gprt_hip.default_config(**head_cfg.as_dict())
    ccfg.use_graph, ccfg.collect_counters = 0, 1
    head_tree.k_nn(dq, K, ccfg, ids_out=ids_b)
    cnt = ctx.stats()
    tcfg = gprt_hip.default_config(**head_cfg.as_dict())
    tcfg.use_graph, tcfg.time_kernels = 0, 1
    kms = []
    for _ in range(max(5, args.steps // 5)):
        head_tree.k_nn(dq, K, tcfg, ids_out=ids_b)
        kms.append(ctx.stats()["ms_kernel"])
    k_ms = float(np.mean(kms))
    r0 = cnt["active_per_round"][0]
    # per query: 16 B query + 64 B per node visit (two child boxes with refs) +
    # 16 B per point test; + 4k B of ids written + 16 B per FP64 re-rank read
    alg_bytes = 16 * r0 + 64 * cnt["node_visits"] + 16 * cnt["point_tests"] + 4 * K * nq + 16 * cnt["rerank_reads"]

    # parity of both forms against the oracle (CPU KD-tree, pinned to brute force)
    head_tree.k_nn(dq, K, head_cfg, ids_out=ids_b)
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
               "seeds": {"lidar_scan": 3}, "traffic_note": traffic_note}
    if auto is not None:
        details["autotune"] = auto
    if not args.no_extra:
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
        steps = max(5, args.steps // 5)
        # k = 32 on the head configuration's tree (the autotuner's pick when
        # adopted: its tree and leaf size) with the default schedule
        ms32, msp32, ib, ip, st = _knn_pair(ctx, head_tree, dq, 32, cfg, steps, 2)
        ref32, _, _, _ = oracle.kdtree_knn(pts, q, 32)
        details["knn_k32"] = {"tree": "sah" if head_sah else "lbvh", "leaf_size": head_leaf,
                              "batched_qps": nq / (ms32 * 1e-3), "plain_qps": nq / (msp32 * 1e-3),
                              "speedup_vs_plain": msp32 / ms32, "ms_batched": ms32, "ms_plain": msp32,
                              "fallback": st["fallback_queries"], "rounds": st["rounds_run"],
                              "batched_eq_oracle": bool(np.array_equal(ib, ref32)),
                              "plain_eq_oracle": bool(np.array_equal(ip, ref32))}
        # the range query on the SAH tree (the counterpart of the reference's
        # default build strategy, PreferFastTrace) with leaf 8: the fastest of
        # leaf 4 / 8 / 16 on either tree for this workload
        rtree = ctx.build(dp, leaf_size=8, sah=True)
        details["range_r0.5"] = _range_line(ctx, rtree, pts, q, oracle)
        details["range_r0.5"].update(tree="sah", leaf_size=8)
        rtree.close()
        details["lbvh_rebuild"] = _build_line(ctx, dp, len(pts), leaf)
        details["sah_build"] = _sah_build_line(ctx, dp, len(pts), head_leaf)
    if args.all:
        details.update(_extra_all(ctx, args, oracle))

    roof = _roof(alg_bytes, k_ms, kernel="knn round 0 traversal (" + ROUND0 + ")",
                 node_visits=cnt["node_visits"], point_tests=cnt["point_tests"], rerank_reads=cnt["rerank_reads"],
                 node_visits_per_s=cnt["node_visits"] / (k_ms * 1e-3))
    roof["traffic"] = traffic["hbm_bytes_per_launch"] if traffic else None
    if traffic:
        roof["traffic_detail"] = traffic
    out = {
        "metric": METRIC,
        "value": nq / (head_ms * 1e-3),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "lidar_scan: 1,310,720 pts (20x64x1024) x 400,000 queries, k=5",
                   "leaf_size": head_leaf, "tree": "sah" if head_sah else "lbvh",
                   "tree_build_ms": head_build_ms, "schedule": head_cfg.as_dict(),
                   "schedule_source": "autotuner" if head_cfg is not cfg else "default",
                   "form": "batched FP32 + FP64 re-rank"},
        "plain": {"value": nq / (ms_p * 1e-3), "ms_per_step": ms_p, "ms_device": ms_p_dev,
                  "form": "one launch, one query per lane in input order, FP64, r=inf"},
        "speedup_vs_plain": ms_p / head_ms,
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity": parity,
        "details": details,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
