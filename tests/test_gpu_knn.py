"""k-NN parity on the GPU: the HIP path (through the C ABI) against the FP64 oracle.

Bar: bit-exact ids (and binary64 distances) against oracle.knn_brute /
oracle.kdtree_knn, which are pinned in test_oracle.py.
"""
import numpy as np
import pytest

import gprt_hip
import oracle
from gprt_hip import datasets as D

pytestmark = pytest.mark.gpu


def _check_knn(ids, ref, d2=None, d2ref=None, label=""):
    ids = np.asarray(ids, dtype=np.uint32).reshape(ref.shape)
    bad = np.nonzero((ids != ref).any(axis=1))[0]
    assert len(bad) == 0, f"{label}: {len(bad)} queries differ, first {bad[:5]}: got {ids[bad[:3]]} want {ref[bad[:3]]}"
    if d2 is not None:
        assert np.array_equal(np.asarray(d2).reshape(d2ref.shape), d2ref), f"{label}: distances differ"


def test_knn_uniform_plain_and_batched(ctx):
    pts, q = D.uniform_cloud()  # 100K x 10K, k = 5
    ref, d2ref = oracle.knn_brute(pts, q, 5, with_d2=True)
    t = ctx.build(pts, leaf_size=8)
    ids, d2 = t.k_nn_plain(q, 5, with_d2=True)
    _check_knn(ids, ref, d2, d2ref, "plain")
    ids, d2 = t.k_nn(q, 5, with_d2=True)
    _check_knn(ids, ref, d2, d2ref, "batched")
    st = ctx.stats()
    assert st["queries"] == len(q)


@pytest.mark.parametrize("k", [1, 2, 5, 13, 16, 29, 32, 37, 38, 45, 61, 62, 64])
def test_knn_k_values(ctx, k):
    rng = np.random.default_rng(10 + k)
    pts = rng.random((20_000, 3), dtype=np.float32)
    q = rng.random((3_000, 3), dtype=np.float32)
    ref = oracle.knn_brute(pts, q, k)
    t = ctx.build(pts, leaf_size=8)
    _check_knn(t.k_nn(q, k), ref, label=f"batched k={k}")
    _check_knn(t.k_nn_plain(q, k), ref, label=f"plain k={k}")
    # the 4-wide walk, also with most queries reaching the unbounded last round
    for kw in (dict(traversal=gprt_hip.TRAV_WIDE),
               dict(traversal=gprt_hip.TRAV_WIDE, radius_heuristic=gprt_hip.RADIUS_FIXED, radius_param=1e-4,
                    max_rounds=1)):
        _check_knn(t.k_nn(q, k, gprt_hip.default_config(**kw)), ref, label=f"batched k={k} {kw}")


@pytest.mark.parametrize("leaf", [1, 2, 4, 16, 32])
def test_knn_leaf_sizes(ctx, leaf):
    rng = np.random.default_rng(leaf)
    pts = rng.normal(size=(30_000, 3)).astype(np.float32)
    q = rng.normal(size=(2_000, 3)).astype(np.float32)
    ref = oracle.knn_brute(pts, q, 5)
    t = ctx.build(pts, leaf_size=leaf)
    _check_knn(t.k_nn(q, 5), ref, label=f"leaf={leaf}")
    _check_knn(t.k_nn_plain(q, 5), ref, label=f"plain leaf={leaf}")


@pytest.mark.parametrize("cfgkw", [
    dict(),
    dict(radius_heuristic=gprt_hip.RADIUS_INFINITE),
    dict(radius_heuristic=gprt_hip.RADIUS_FIXED, radius_param=1e-4, max_rounds=15, radius_mult=2.0),
    dict(radius_heuristic=gprt_hip.RADIUS_FIXED, radius_param=1e-4, max_rounds=3),
    dict(sampler=gprt_hip.SAMPLER_STRIDED, radius_param=0.10),
    dict(radius_heuristic=gprt_hip.RADIUS_MAX),
    dict(sort_queries=0),
    dict(use_graph=0),
    dict(stack_depth=32, grid_blocks=7),
    dict(lds_stack=8, stack_depth=24),   # most walks spill to HBM
    dict(lds_stack=8, stack_depth=9),    # walks overflow: exact fallback
    dict(lds_stack=32, stack_depth=32),  # no spill
    dict(traversal=gprt_hip.TRAV_WIDE),
    dict(traversal=gprt_hip.TRAV_WIDE, lds_stack=8, stack_depth=10),  # wide walks overflow: exact fallback
    dict(traversal=gprt_hip.TRAV_WIDE, radius_heuristic=gprt_hip.RADIUS_INFINITE, lds_stack=16),
    dict(traversal=gprt_hip.TRAV_WAVE),
    dict(traversal=gprt_hip.TRAV_WIDE, radius_heuristic=gprt_hip.RADIUS_FIXED, radius_param=1e-4, max_rounds=2),
    dict(compact_rounds=1),
    dict(compact_rounds=1, radius_heuristic=gprt_hip.RADIUS_FIXED, radius_param=1e-4, max_rounds=15, radius_mult=2.0),
    dict(traversal=gprt_hip.TRAV_WIDE, radius_heuristic=gprt_hip.RADIUS_FIXED, radius_param=1e-4, max_rounds=15,
         radius_mult=2.0),
])
def test_knn_config_independence(ctx, cfgkw):
    """The result must not depend on the initial radius, multiplier, rounds,
    query order, graph use, stack depth and LDS ring size, or grid size."""
    pts, q = D.uniform_cloud(50_000, 5_000, seed=7)
    ref = oracle.knn_brute(pts, q, 5)
    t = ctx.build(pts)
    cfg = gprt_hip.default_config(**cfgkw)
    _check_knn(t.k_nn(q, 5, cfg), ref, label=str(cfgkw))
    # second call replays the cached graph
    _check_knn(t.k_nn(q, 5, cfg), ref, label=str(cfgkw) + " replay")


def test_knn_ties_lattice(ctx):
    """Integer lattice: many exactly equal distances; ids break the ties and the
    guard band overflows into the exact kernel."""
    g = np.arange(24, dtype=np.float32)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).copy()
    rng = np.random.default_rng(1)
    q = np.concatenate([pts[rng.integers(0, len(pts), 500)],
                        pts[rng.integers(0, len(pts), 500)] + 0.5,
                        rng.integers(0, 24, (500, 3)).astype(np.float32) + np.float32(0.25)]).astype(np.float32)
    for k in (1, 6, 7, 27):
        ref = oracle.knn_brute(pts, q, k)
        t = ctx.build(pts, leaf_size=4)
        _check_knn(t.k_nn(q, k), ref, label=f"lattice k={k}")
        _check_knn(t.k_nn_plain(q, k), ref, label=f"lattice plain k={k}")
    assert ctx.stats()["queries"] == len(q)


@pytest.mark.parametrize("case", ["n1437", "n5000"])
def test_knn_lattice_ties_large_k(ctx, case):
    """Cases a randomized 400-case sweep found: an integer lattice of 1437 points
    with k = 55 (rounds retried inline with radius_mult 1.1), and the 5000-point
    lattice of the parity matrix at k = 60, 61 and 63.  Both forms must equal the
    oracle."""
    from test_gpu_parity_matrix import _cloud, _queries
    if case == "n1437":
        rng = np.random.default_rng(90063)
        rng.integers(6), rng.uniform(0, np.log(100_000)), rng.integers(1, 65), rng.choice(6), rng.integers(10)
        pts = _cloud("lattice", 1437, rng)
        q = _queries(pts, 500, rng)
        t, ks, cfg = ctx.build(pts, leaf_size=32), (55,), gprt_hip.default_config(radius_mult=1.1, max_rounds=3)
    else:
        rng = np.random.default_rng(1064)
        pts = _cloud("lattice", 5000, rng)
        q = _queries(pts, 1500, rng)
        t, ks, cfg = ctx.build(pts, leaf_size=8), (60, 61, 63), gprt_hip.default_config()
    try:
        for k in ks:
            ref = oracle.knn_brute(pts, q, k)
            _check_knn(t.k_nn_plain(q, k), ref, label=f"lattice plain k={k}")
            _check_knn(t.k_nn(q, k, cfg), ref, label=f"lattice k={k}")
    finally:
        t.close()


@pytest.mark.parametrize("k", [46, 64, 65, 100, 256, 1000, 4096])
def test_knn_large_k(ctx, k):
    """k beyond the per-lane list (k + 3 > 48): one query per wavefront with the
    list spread over its lanes, in both forms, on a uniform cloud and on an
    integer lattice full of exact distance ties.  Clouds of fewer than k points
    pad with 0xFFFFFFFF."""
    rng = np.random.default_rng(700 + k)
    pts = rng.random((20_000, 3), dtype=np.float32)
    q = np.concatenate([rng.random((300, 3), dtype=np.float32), pts[:100]])
    g = np.arange(20, dtype=np.float32)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).copy()
    ql = np.concatenate([lat[rng.integers(0, len(lat), 200)],
                         lat[rng.integers(0, len(lat), 100)] + np.float32(0.5)]).astype(np.float32)
    small = rng.random((k // 2 + 1, 3), dtype=np.float32)
    for label, P, Q in (("uniform", pts, q), ("lattice", lat, ql), ("n<k", small, q[:50])):
        ref = oracle.knn_brute(P, Q, k)
        t = ctx.build(P, leaf_size=8)
        try:
            _check_knn(t.k_nn(Q, k), ref, label=f"{label} batched k={k}")
            _check_knn(t.k_nn_plain(Q, k), ref, label=f"{label} plain k={k}")
            if k <= 256:
                cfg = gprt_hip.default_config(radius_heuristic=gprt_hip.RADIUS_FIXED, radius_param=1e-3,
                                              max_rounds=6, radius_mult=2.0)
                _check_knn(t.k_nn(Q, k, cfg), ref, label=f"{label} tiny radius k={k}")
        finally:
            t.close()


def test_knn_guard_slots_fixed(ctx):
    """The lists carry exactly three guard slots; a schedule asking for another
    number is rejected rather than ignored."""
    pts = np.random.default_rng(0).random((100, 3), dtype=np.float32)
    t = ctx.build(pts)
    for g in (0, 1, 2, 4):
        with pytest.raises(gprt_hip.GprtError):
            t.k_nn(pts, 5, gprt_hip.default_config(guard_slots=g))
    t.close()


def test_knn_k_limit(ctx):
    pts = np.random.default_rng(0).random((100, 3), dtype=np.float32)
    t = ctx.build(pts)
    with pytest.raises(gprt_hip.GprtError):
        t.k_nn(pts, gprt_hip.K_MAX + 1)
    with pytest.raises(gprt_hip.GprtError):
        t.k_nn_plain(pts, gprt_hip.K_MAX + 1)
    t.close()


def test_knn_duplicates(ctx):
    rng = np.random.default_rng(3)
    base = rng.random((200, 3), dtype=np.float32)
    pts = np.repeat(base, 50, axis=0)  # every point 50 times
    q = base[rng.integers(0, 200, 1000)]
    ref = oracle.knn_brute(pts, q, 8)
    t = ctx.build(pts)
    _check_knn(t.k_nn(q, 8), ref, label="duplicates")
    _check_knn(t.k_nn_plain(q, 8), ref, label="duplicates plain")


@pytest.mark.parametrize("n", [1, 2, 3, 5, 9])
def test_knn_small_clouds(ctx, n):
    """Clouds smaller than k: n ids then 0xFFFFFFFF padding."""
    rng = np.random.default_rng(n)
    pts = rng.random((n, 3), dtype=np.float32)
    q = rng.random((100, 3), dtype=np.float32)
    ref = oracle.knn_brute(pts, q, 5)
    t = ctx.build(pts)
    _check_knn(t.k_nn(q, 5), ref, label=f"n={n}")
    _check_knn(t.k_nn_plain(q, 5), ref, label=f"plain n={n}")


def test_knn_empty_inputs(ctx):
    pts = np.random.default_rng(0).random((1000, 3), dtype=np.float32)
    t = ctx.build(pts)
    assert t.k_nn(np.zeros((0, 3), np.float32), 5).shape == (0, 5)
    t0 = ctx.build(np.zeros((0, 3), np.float32))
    ids = t0.k_nn(pts[:10], 3)
    assert (ids == gprt_hip.ID_NONE).all()


def test_knn_rejects_nan(ctx):
    pts = np.random.default_rng(0).random((1000, 3), dtype=np.float32)
    bad = pts.copy()
    bad[17, 1] = np.nan
    with pytest.raises(gprt_hip.GprtError) as e:
        ctx.build(bad)
    assert e.value.code == gprt_hip.ERR_INPUT
    t = ctx.build(pts)
    q = pts[:100].copy()
    q[3, 0] = np.inf
    with pytest.raises(gprt_hip.GprtError) as e:
        t.k_nn(q, 5)
    assert e.value.code == gprt_hip.ERR_INPUT


def test_knn_road_network_subset(ctx):
    """1e5-metre coordinates: FP32 spacing ~8 mm, many near-ties in FP32."""
    pts, q = D.road_network(nq=50_000)
    ref, _, _, _ = oracle.kdtree_knn(pts, q, 5)
    t = ctx.build(pts)
    _check_knn(t.k_nn(q, 5), ref, label="road batched")
    _check_knn(t.k_nn_plain(q, 5), ref, label="road plain")


def test_knn_lidar_full(ctx):
    """The headline workload at full size: 1,310,720 points x 400,000 queries,
    k = 5 and k = 32, against the CPU KD-tree (itself pinned to brute force);
    k = 5 also with an 8-entry LDS ring (deep walks spill to HBM) and with
    one unbounded closest-first round."""
    pts, q = D.lidar_scan()
    t = ctx.build(pts)
    for k in (5, 32):
        ref, _, _, _ = oracle.kdtree_knn(pts, q, k)
        _check_knn(t.k_nn(q, k), ref, label=f"lidar batched k={k}")
        st = ctx.stats()
        print(f"lidar k={k}: rounds={st['rounds_run']} fallback={st['fallback_queries']} "
              f"active={st['active_per_round'][:8]} ms={st['ms_query']:.3f}")
        _check_knn(t.k_nn_plain(q, k), ref, label=f"lidar plain k={k}")
        if k == 5:
            for kw in (dict(lds_stack=8), dict(lds_stack=8, radius_heuristic=gprt_hip.RADIUS_INFINITE),
                       dict(traversal=gprt_hip.TRAV_WIDE), dict(traversal=gprt_hip.TRAV_WIDE, lds_stack=16)):
                _check_knn(t.k_nn(q, k, gprt_hip.default_config(**kw)), ref, label=f"lidar batched {kw}")
                assert ctx.stats()["fallback_queries"] == 0


def test_knn_device_pointers(ctx):
    """Inputs and outputs resident in HBM (GPRT_PTR_DEVICE)."""
    pts, q = D.uniform_cloud(50_000, 4_000, seed=9)
    ref = oracle.knn_brute(pts, q, 5)
    dp = gprt_hip.DeviceArray.from_host(ctx, pts)
    dq = gprt_hip.DeviceArray.from_host(ctx, q)
    t = ctx.build(dp)
    _check_knn(t.k_nn(dq, 5).to_host(), ref, label="device batched")
    _check_knn(t.k_nn_plain(dq, 5).to_host(), ref, label="device plain")
    # enqueue-only calls, then one sync
    out = gprt_hip.DeviceArray(ctx, (len(q), 5), np.uint32)
    for _ in range(3):
        t.k_nn(dq, 5, ids_out=out, sync=False)
    ctx.sync()
    _check_knn(out.to_host(), ref, label="device no-sync")


@pytest.mark.parametrize("bits", [8, 12, 16, 20])
def test_query_cell_sort_bits(ctx, bits):
    """Query ordering by 8 to 20 Morton bits (one to three radix passes):
    results are unchanged."""
    L = gprt_hip.lib()
    pts, q = D.uniform_cloud(50_000, 5_000, seed=8)
    ref = oracle.knn_brute(pts, q, 5)
    t = ctx.build(pts)
    L.gprt_hip_debug_set_query_sort_bits(bits)
    try:
        cfg = gprt_hip.default_config(use_graph=0)
        for _ in range(2):
            _check_knn(t.k_nn(q, 5, cfg), ref, label=f"bits={bits}")
        off, ids = t.range_query(q[:800], 0.03)
        roff, rids = oracle.range_brute(pts, q[:800], 0.03)
        assert np.array_equal(off, roff) and np.array_equal(ids, rids)
    finally:
        L.gprt_hip_debug_set_query_sort_bits(16)
