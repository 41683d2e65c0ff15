"""k-NN command-line harness with the reference sandbox's arguments and output.

    python -m gprt_hip.sandbox <data.csv> <queries.csv> <k> [percentile] [mult] [dump] [tune]

Follows gprt-sandbox/src/main.rs:10-97:

- arguments: after k, "dump" and "tune" are flags and the first two numbers
  are the percentile (default 0.10) and the radius multiplier (default 3.0);
- "[IO] Loaded N points and M queries in X.XXms" after reading both files
  (gprt_hip.read_points, the rules of gprt-sandbox/src/io.rs:8-30);
- "tune": runs the autotuner over the full inputs and saves gprt_wisdom.json in
  the working directory; otherwise gprt_wisdom.json there is loaded when it
  reads as a schedule ("[LOAD] Loaded auto-tuned schedule from
  gprt_wisdom.json"), a file that does not read falls back to the command line
  silently, and with no file "[LOAD] Using CLI defaults: P=0.10, M=3.0";
- the timed region is the whole k_nn call: tree build, radius sampling, the
  rounds, and the copies of queries in and ids out;
- "[GPRT_LOOP]" lines give the sampled radius and the queries entering each
  round; "[GPRT_STATS] search_ms=..., saturation=...%" gives the query phase's
  device time and the share of queries finished by the exact FP64 kernel (the
  counterpart of a query whose hit buffer saturated);
- "dump" writes knn_results.txt, one "Q{i}:[id, id, ...]" line per query in
  input order, ids nearest first (fewer than k when the data holds fewer);
- last line: "{data_path},{queries},{k},{total_ms:.3},{neighbours}".

Unlike the reference, a wisdom file that lacks the Schedule fields added after
it was written (build_strategy, geom_type) still loads: missing fields take
Schedule::default().
"""
import os
import sys
import time

import numpy as np

import gprt_hip

WISDOM = "gprt_wisdom.json"


def parse_args(argv):
    """(data_path, query_path, k, percentile, mult, dump, tune) as main.rs:12-38
    reads them; raises SystemExit(1) with the usage line on too few arguments."""
    if len(argv) < 4:
        print(f"Usage: {argv[0] if argv else 'sandbox'} <data.csv> <queries.csv> <k> [percentile] [mult] [dump] [tune]",
              file=sys.stderr)
        raise SystemExit(1)
    try:
        k = int(argv[3])
    except ValueError:
        raise SystemExit("K must be an integer")
    percentile, mult, dump, tune = 0.10, 3.0, False, False
    nums = 0
    for a in argv[4:]:
        if a == "dump":
            dump = True
        elif a == "tune":
            tune = True
        else:
            try:
                v = float(np.float32(float(a)))
            except ValueError:
                continue
            if nums == 0:
                percentile = v
            elif nums == 1:
                mult = v
            nums += 1
    return argv[1], argv[2], k, percentile, mult, dump, tune


def choose_schedule(ctx, data, queries, k, percentile, mult, tune, out=print):
    """(Config, leaf_size, sah) per main.rs:50-76.  The tree follows the
    schedule's build_strategy (PreferFastTrace: the SAH build): the autotuner
    chooses it, a loaded file names it (Schedule::default() when absent); the
    command-line defaults build the LBVH, whose build is the faster one inside
    the timed region."""
    if tune:
        out("[TUNE] Starting Native Full-Dataset Autotuner...")
        cfg, leaf, sah = ctx.autotune(data, queries, k, json_path=WISDOM, with_tree=True)
        out(f"[TUNE] Saved optimal schedule to {WISDOM}")
        return cfg, leaf, sah
    cfg, leaf = gprt_hip.parse_schedule("{}")  # Schedule::default()
    if os.path.exists(WISDOM):
        try:
            cfg, leaf, sah = gprt_hip.load_schedule(WISDOM, with_tree=True)
            out(f"[LOAD] Loaded auto-tuned schedule from {WISDOM}")
            return cfg, leaf, sah
        except (gprt_hip.GprtError, OSError, UnicodeDecodeError):
            pass
    else:
        out(f"[LOAD] Using CLI defaults: P={percentile:.2f}, M={mult:.1f}")
    cfg.radius_heuristic = gprt_hip.RADIUS_PERCENTILE
    cfg.radius_param = percentile
    cfg.radius_mult = mult
    return cfg, leaf, False


def format_results(ids):
    """knn_results.txt text: "Q{i}:[a, b, ...]" per query, padding dropped."""
    lines = []
    for i, row in enumerate(ids):
        row = row[row != gprt_hip.ID_NONE]
        lines.append(f"Q{i}:[{', '.join(map(str, row.tolist()))}]")
    return "\n".join(lines) + ("\n" if lines else "")


def run(argv, out=print):
    data_path, query_path, k, percentile, mult, dump, tune = parse_args(argv)
    t0 = time.perf_counter()
    data = gprt_hip.read_points(data_path)
    queries = gprt_hip.read_points(query_path)
    out(f"[IO] Loaded {len(data)} points and {len(queries)} queries in {(time.perf_counter() - t0) * 1e3:.2f}ms")
    ctx = gprt_hip.Context(0)
    try:
        cfg, leaf, sah = choose_schedule(ctx, data, queries, k, percentile, mult, tune, out)
        t1 = time.perf_counter()
        tree = ctx.build(data, leaf_size=leaf, sah=sah)
        ids = tree.k_nn(queries, k, cfg)
        total_ms = (time.perf_counter() - t1) * 1e3
        st = ctx.stats()
        tree.close()
    finally:
        ctx.close()
    out(f"[GPRT_LOOP] Initial Radius Calculated: {st['initial_radius']:.5f}")
    r = st["initial_radius"]
    for i, active in enumerate(st["active_per_round"][:max(st["rounds_run"], 1)]):
        out(f"[GPRT_LOOP] Round {i}: Active={active}, Radius={r:.5f}")
        r *= cfg.radius_mult
    sat = st["fallback_queries"] / len(queries) if len(queries) else 0.0
    out(f"[GPRT_STATS] search_ms={st['ms_query']:.3f}, saturation={sat * 100:.2f}%")
    total = int((ids != gprt_hip.ID_NONE).sum())
    if dump:
        with open("knn_results.txt", "w") as f:
            f.write(format_results(ids))
    out(f"{data_path},{len(queries)},{k},{total_ms:.3f},{total}")
    return ids


def main(argv=None):
    run(sys.argv if argv is None else argv)


if __name__ == "__main__":
    main()
