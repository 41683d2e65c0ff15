"""GPU parity of the building blocks and the other hot-path entry points:
Morton codes, radius samplers, radix sort, LBVH invariants, range query,
batched uneven clouds, Barnes-Hut, the reference FFI, the autotuner."""
import ctypes
import json
import os

import numpy as np
import pytest

import gprt_hip
import oracle
from gprt_hip import datasets as D
from gprt_hip import hostwalk as H
from test_host_bh import EPS, F64_SLACK, ROUND_U, U, bound_violations, rounding_violations

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------- pieces
@pytest.mark.parametrize("case", ["uniform", "flat", "negative", "road", "lidar_subset"])
def test_morton_codes_bit_exact(ctx, case):
    rng = np.random.default_rng(1)
    if case == "uniform":
        pts = rng.random((100_000, 3), dtype=np.float32)
    elif case == "flat":
        pts = rng.random((10_000, 3), dtype=np.float32)
        pts[:, 2] = 7.0
    elif case == "negative":
        pts = (rng.normal(size=(10_000, 3)) * 50).astype(np.float32)
    elif case == "road":
        pts, _ = D.road_network(n=50_000, nq=10)
    else:
        pts, _ = D.lidar_scan(frames=2, nq=10)
    codes, bbox = ctx.morton_codes(pts)
    rc, rb = oracle.morton_codes(pts)
    assert np.array_equal(bbox, rb)
    assert np.array_equal(codes, rc)


@pytest.mark.parametrize("sampler,heur,param,k", [
    (gprt_hip.SAMPLER_LCG, gprt_hip.RADIUS_PERCENTILE, 0.95, 5),
    (gprt_hip.SAMPLER_LCG, gprt_hip.RADIUS_PERCENTILE, 0.10, 5),
    (gprt_hip.SAMPLER_LCG, gprt_hip.RADIUS_MAX, 0.0, 5),
    (gprt_hip.SAMPLER_LCG, gprt_hip.RADIUS_FIXED, 0.25, 5),
    (gprt_hip.SAMPLER_STRIDED, gprt_hip.RADIUS_PERCENTILE, 0.10, 5),
    (gprt_hip.SAMPLER_STRIDED, gprt_hip.RADIUS_PERCENTILE, 0.95, 16),
    (gprt_hip.SAMPLER_STRIDED, gprt_hip.RADIUS_FIXED, 0.5, 5),
])
def test_samplers_bit_exact(ctx, sampler, heur, param, k):
    for pts, q in (D.uniform_cloud(100_000, 10_000), D.lidar_scan(frames=3, nq=20_000)):
        got = ctx.sample_radius(pts, q, sampler, heur, param, k)
        if sampler == gprt_hip.SAMPLER_LCG:
            want = oracle.sampler_lcg(pts, q, heur, param)
        else:
            want = oracle.sampler_strided(pts, q, k, heur, param)
        assert got == want, (got, want)


@pytest.mark.parametrize("heur,param", [(gprt_hip.RADIUS_PERCENTILE, 0.95), (gprt_hip.RADIUS_PERCENTILE, 0.1),
                                         (gprt_hip.RADIUS_MAX, 0.0)])
def test_knn_uses_sampled_radius(ctx, heur, param):
    """The radius a k-NN call starts from (run statistics) is the standalone
    sampler's and the oracle's, eager and on replays of the captured graph."""
    pts, q = D.lidar_scan(frames=3, nq=20_000)
    want = oracle.sampler_lcg(pts, q, heur, param)
    assert ctx.sample_radius(pts, q, gprt_hip.SAMPLER_LCG, heur, param, 5) == want
    t = ctx.build(pts)
    ref = oracle.knn_brute(pts, q, 5)
    for use_graph in (0, 1):
        cfg = gprt_hip.default_config(radius_heuristic=heur, radius_param=param, use_graph=use_graph)
        for _ in range(3):
            assert np.array_equal(t.k_nn(q, 5, cfg), ref)
            assert np.float32(ctx.stats()["initial_radius"]) == np.float32(want)


@pytest.mark.parametrize("n,nq", [(1, 1), (2, 3), (7, 1000), (4_999, 99), (5_000, 100), (5_001, 101),
                                  (65_537, 7), (262_147, 50_000), (1_048_573, 123_457)])
def test_sampler_lcg_sizes(ctx, n, nq):
    """The LCG sampler's draws (lane jumps from the compile-time table, indices
    reduced mod n without integer division) across cloud and query sizes,
    prime and power-of-two neighbours included, bit for bit."""
    rng = np.random.default_rng(n + nq)
    pts = rng.random((n, 3), dtype=np.float32)
    q = rng.random((nq, 3), dtype=np.float32)
    for heur, param in ((gprt_hip.RADIUS_PERCENTILE, 0.95), (gprt_hip.RADIUS_MAX, 0.0)):
        assert ctx.sample_radius(pts, q, gprt_hip.SAMPLER_LCG, heur, param, 5) == \
            oracle.sampler_lcg(pts, q, heur, param)


def test_sampler_small_inputs(ctx):
    pts = np.zeros((3, 3), np.float32)
    q = np.ones((5, 3), np.float32)
    assert ctx.sample_radius(pts, q, gprt_hip.SAMPLER_LCG, gprt_hip.RADIUS_PERCENTILE, 0.5) == 10.0
    assert ctx.sample_radius(pts, q, gprt_hip.SAMPLER_LCG, gprt_hip.RADIUS_PERCENTILE, 0.5) == \
        oracle.sampler_lcg(pts, q, oracle.HEUR_PERCENTILE, 0.5)


@pytest.mark.parametrize("n,bits", [(1, (0, 64)), (1000, (0, 30)), (300_000, (0, 40)), (70_000, (8, 24))])
def test_radix_sort_stable(ctx, n, bits):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 40, n, dtype=np.uint64) & np.uint64((1 << 40) - 1)
    keys[: n // 3] = keys[0]  # many duplicates
    vals = np.arange(n, dtype=np.uint32)
    k2, v2 = ctx.sort_pairs(keys, vals, *bits)
    sub = (keys >> np.uint64(bits[0])) & np.uint64((1 << (bits[1] - bits[0])) - 1) if bits[1] < 64 else keys >> np.uint64(bits[0])
    order = np.argsort(sub, kind="stable")
    assert np.array_equal(v2, vals[order])
    assert np.array_equal(k2, keys[order])


def _walk(nodes, roots, c):
    """Host walk of the exported tree: every point of cloud c once; boxes tight."""
    ref, cnt = roots[c]
    seen = []
    if cnt == 0:
        return seen
    if ref < 0:
        return list(range(~ref, ~ref + cnt))
    stack = [ref]
    while stack:
        i = stack.pop()
        nd = nodes[i]
        for lo, hi in ((0, 4), (8, 12)):
            r = nd[lo + 3].view(np.int32)
            m = nd[hi + 3].view(np.int32)
            if r < 0:
                seen.extend(range(~r, ~r + m))
            else:
                stack.append(int(r))
    return seen


@pytest.mark.parametrize("leaf,ref_keys", [(1, False), (4, False), (8, False), (32, False), (4, True), (8, True)])
def test_lbvh_invariants(ctx, leaf, ref_keys):
    pts, _ = D.road_network(n=40_000, nq=10)
    t = ctx.build(pts, leaf_size=leaf, ref_keys=ref_keys)
    info = t.info()
    bits = 10 if ref_keys else 21
    assert info["axis_bits"] == bits and info["key_bits"] == 3 * bits
    keys, perm, nodes, roots = t.export()
    assert np.array_equal(np.sort(perm), np.arange(len(pts)))
    assert np.all(keys[1:] >= keys[:-1])
    # the tree keys are Morton codes over a cube of the largest extent
    mn, mx = pts.min(0), pts.max(0)
    e = np.float32((mx - mn).max())
    kbox = np.concatenate([mn, (mn + e).astype(np.float32)]).astype(np.float32)
    codes = oracle.morton_codes_wide(pts, kbox, bits)
    assert np.array_equal(keys, codes[perm])
    if ref_keys:
        assert np.array_equal(keys.astype(np.uint32), oracle.morton_codes_bbox(pts, kbox)[perm])
    # stable order: equal codes keep input order
    same = keys[1:] == keys[:-1]
    assert np.all(np.diff(perm.astype(np.int64))[same] > 0)
    seen = _walk(nodes, roots, 0)
    assert sorted(seen) == list(range(len(pts)))
    # child boxes equal the union of their points (sorted order)
    sp = pts[perm]
    rng = np.random.default_rng(0)
    for i in rng.integers(0, len(nodes), 300):
        nd = nodes[i]
        for lo, hi in ((0, 4), (8, 12)):
            r = nd[lo + 3].view(np.int32)
            m = nd[hi + 3].view(np.int32)
            if r < 0:
                sub = sp[~r: ~r + m]
                assert np.array_equal(nd[lo:lo + 3], sub.min(0)) and np.array_equal(nd[hi:hi + 3], sub.max(0))


def test_lbvh_batched_roots(ctx):
    pts, poff, q, qoff, kc = D.clustered_batch(n_clouds=12, lo=1, hi=5e3, seed=3)
    t = ctx.build_batched(pts, poff, leaf_size=8)
    keys, perm, nodes, roots = t.export()
    for c in range(12):
        got = sorted(int(perm[i]) for i in _walk(nodes, roots, c))
        assert got == list(range(poff[c], poff[c + 1])), c


# ---------------------------------------------------------------- range query
def _check_range(off, ids, roff, rids):
    assert np.array_equal(off, roff)
    assert np.array_equal(ids, rids)


def test_range_uniform(ctx):
    pts, q = D.uniform_cloud(50_000, 3_000, seed=4)
    t = ctx.build(pts)
    for r in (0.0, 0.01, 0.05, 0.2):
        _check_range(*t.range_query(q, r), *oracle.range_brute(pts, q, r))


def test_range_boundary_and_lattice(ctx):
    g = np.arange(12, dtype=np.float32)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).copy()
    q = pts[::7].copy()
    t = ctx.build(pts, leaf_size=4)
    for r in (1.0, np.sqrt(2.0), 2.0, 3.0):  # many points exactly on the sphere
        _check_range(*t.range_query(q, float(np.float32(r))), *oracle.range_brute(pts, q, float(np.float32(r))))


def test_range_long_segments(ctx):
    """Result lists longer than one LDS sort (2,048) and several merge levels:
    a dense ball around the queries, lists of ~3K to ~15K ids."""
    rng = np.random.default_rng(12)
    ball = rng.normal(0.0, 0.05, (15_000, 3)).astype(np.float32)
    rest = rng.random((30_000, 3)).astype(np.float32) * 4 - 2
    pts = np.concatenate([ball, rest])
    q = np.concatenate([ball[:40], rest[:40], np.zeros((1, 3), np.float32)])
    t = ctx.build(pts, leaf_size=8)
    for r in (0.06, 0.3):
        off, ids = t.range_query(q, r)
        assert np.diff(off).max() > 2048
        _check_range(off, ids, *oracle.range_brute(pts, q, r))


def test_range_segment_tiers(ctx):
    """Every sort tier at its edges: points at the integers of a line, query i at
    x = 20000.5 with r = m/2 hits exactly m points, in pairs at equal distance
    (ties broken by id), for m on both sides of 1, 64, 2,048, 8,192 and
    12,288 (sorted in LDS parts up to there) and beyond (chunks plus merges);
    plus a segment of 9,000 coincident points (one bucket of more than 8,192:
    left to the chunks and merges of all-equal distances)."""
    x = np.arange(40_001, dtype=np.float32)
    line = np.c_[x, np.zeros((len(x), 2), np.float32)].astype(np.float32)
    dup = np.tile(np.float32([[5e4, 7.0, 7.0]]), (9_000, 1))
    pts = np.concatenate([line, dup])
    t = ctx.build(pts, leaf_size=8)
    try:
        for m in (2, 64, 66, 2048, 2050, 8192, 8194, 12_288, 12_290, 17_000):
            q = np.float32([[20000.5, 0.0, 0.0]])
            r = float(np.float32(m / 2))
            off, ids = t.range_query(q, r)
            assert off[-1] == m, (m, off)
            _check_range(off, ids, *oracle.range_brute(pts, q, r))
        q = np.float32([[20000.0, 0.0, 0.0]])  # one hit
        off, ids = t.range_query(q, 0.25)
        assert off[-1] == 1
        _check_range(off, ids, *oracle.range_brute(pts, q, 0.25))
        q = np.float32([[5e4, 7.0, 7.0], [5e4, 7.0, 7.5]])
        for r in (0.0, 0.5):
            _check_range(*t.range_query(q, r), *oracle.range_brute(pts, q, r))
    finally:
        t.close()


def test_range_whole_subtrees(ctx):
    """Subtrees inside the ball are emitted as runs and expanded by the sorts:
    a radius covering the whole cloud (the root's two children taken whole, one
    run of every point per query, lists in the chunked-merge tier), a dense
    lattice where most hits come from whole subtrees, radii so small that the
    inner threshold proves nothing, and points on the ball's surface."""
    rng = np.random.default_rng(31)
    pts = rng.random((12_000, 3)).astype(np.float32)
    q = np.concatenate([rng.random((6, 3)), [[0.5, 0.5, 0.5], [3.0, 3.0, 3.0]]]).astype(np.float32)
    t = ctx.build(pts, leaf_size=8)
    try:
        for r in (2.0, 5.0):  # every point of the cloud for the queries in the unit cube
            off, ids = t.range_query(q[:7], r)
            assert np.all(np.diff(off) == len(pts))
            _check_range(off, ids, *oracle.range_brute(pts, q[:7], r))
        for r in (1e-20, 1e-19, 0.0):
            _check_range(*t.range_query(q, r), *oracle.range_brute(pts, q, r))
    finally:
        t.close()
    g = np.arange(40, dtype=np.float32) * np.float32(0.25)
    lat = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).copy()
    ql = np.concatenate([lat[::997], lat[::1311] + np.float32(0.125)]).astype(np.float32)
    t = ctx.build(lat, leaf_size=4)
    try:
        for r in (0.5, 1.0, 1.75, 2.5):  # lattice points exactly on the sphere, long lists
            rr = float(np.float32(r))
            off, ids = t.range_query(ql, rr, stats=(st := []))
            assert st[1]["rerank_reads"] == 0  # counters off
            _check_range(off, ids, *oracle.range_brute(lat, ql, rr))
        st = []
        off, ids = t.range_query(ql, 1.75, flags=gprt_hip.COLLECT_COUNTERS, stats=st)
        # slots written (positions and two-slot runs) are far fewer than hits
        assert 0 < st[1]["rerank_reads"] < off[-1] // 2, (st[1]["rerank_reads"], off[-1])
        assert st[0]["rerank_reads"] == st[1]["rerank_reads"]
    finally:
        t.close()


def test_range_fill_offsets_mismatch(ctx):
    """A fill call whose offsets do not match the queries and radius fails with
    GPRT_ERR_ARG (the walk checks every range it fills) and writes nowhere
    outside the ids array; the context keeps working."""
    pts, q = D.uniform_cloud(20_000, 500, seed=8)
    t = ctx.build(pts)
    L = gprt_hip.lib()
    try:
        off, ids = t.range_query(q, 0.08)
        qq = np.ascontiguousarray(q, np.float32)
        for bad_off in (off // 2, np.concatenate([[0], np.full(len(q), off[-1] // 2)]).astype(np.int64)):
            bad_off = np.ascontiguousarray(bad_off, np.int64)
            total = int(bad_off[-1])
            buf = np.full(total + 64, 0xA5A5A5A5, np.uint32)
            rc = L.gprt_hip_range(ctx.h, t.h, qq.ctypes.data, len(q), ctypes.c_float(0.08), bad_off.ctypes.data,
                                  buf.ctypes.data, total, gprt_hip.PTR_HOST)
            assert rc == 1, rc  # GPRT_ERR_ARG
            assert "offsets_out does not match" in L.gprt_hip_last_error().decode()
            assert np.all(buf[total:] == 0xA5A5A5A5)
        _check_range(*t.range_query(q, 0.08), off, ids)
    finally:
        t.close()


def test_range_lidar_subset(ctx):
    pts, q = D.lidar_scan(frames=4, nq=20_000)
    t = ctx.build(pts)
    off, ids = t.range_query(q, 0.5)
    roff, rids = oracle.range_brute(pts, q, 0.5)
    _check_range(off, ids, roff, rids)
    assert off[-1] > len(q)  # non-trivial result sizes, including long lists


# ---------------------------------------------------------------- batched uneven clouds
def test_knn_batched_uneven(ctx):
    pts, poff, q, qoff, kc = D.clustered_batch(n_clouds=24, lo=1e2, hi=2e4, seed=6, ks=(1, 5, 16, 32, 64))
    t = ctx.build_batched(pts, poff)
    ids, out_off = t.k_nn_batched(q, qoff, kc)
    for c in range(24):
        P = pts[poff[c]:poff[c + 1]]
        Q = q[qoff[c]:qoff[c + 1]]
        if len(Q) == 0:
            continue
        ref = oracle.knn_brute(P, Q, int(kc[c]))
        got = ids[out_off[c]:out_off[c + 1]].reshape(len(Q), int(kc[c]))
        assert np.array_equal(got, ref), c


def test_knn_batched_tiny_and_empty_clouds(ctx):
    rng = np.random.default_rng(2)
    sizes = [0, 1, 3, 200, 0, 5000]
    pts = rng.random((sum(sizes), 3), dtype=np.float32)
    poff = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    qs = [rng.random((n_, 3), dtype=np.float32) for n_ in (4, 0, 6, 50, 3, 100)]
    qoff = np.concatenate([[0], np.cumsum([len(x) for x in qs])]).astype(np.int64)
    kc = np.array([5, 5, 5, 8, 2, 13], np.int32)
    t = ctx.build_batched(pts, poff)
    ids, out_off = t.k_nn_batched(np.concatenate(qs), qoff, kc)
    for c in range(6):
        if len(qs[c]) == 0:
            continue
        ref = oracle.knn_brute(pts[poff[c]:poff[c + 1]], qs[c], int(kc[c]))
        assert np.array_equal(ids[out_off[c]:out_off[c + 1]].reshape(len(qs[c]), -1), ref), c


# ---------------------------------------------------------------- Barnes-Hut
def _rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)


def _bh_same_as_host(acc, h):
    return not (acc.view(np.uint32) != h["acc32"].view(np.uint32)).any()


def test_bh_theta_zero_is_direct_sum(ctx):
    """theta = 0 opens every node: the FP64 direct sum to the rounding
    allowance of FP32 terms and compensated sums (test_host_bh.py)."""
    pm = D.plummer(20_000, seed=7)
    acc = ctx.barnes_hut(pm, theta=0.0, eps=EPS, G=1.0)
    sample = np.arange(0, 20_000, 97)
    h = H.barnes_hut(pm, 0.0, eps=EPS, sample=sample)
    assert _bh_same_as_host(acc[sample], h)
    assert len(bound_violations(acc[sample], h, oracle.bh_direct(pm, sample, G=1.0, eps=EPS))) == 0


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_bh_error_vs_direct(ctx, theta):
    """100,000 Plummer bodies: the host build bit for bit and the derived
    monopole bound against the FP64 direct sum; errors printed."""
    pm = D.plummer(100_000, seed=8)
    acc = ctx.barnes_hut(pm, theta=theta, eps=EPS, G=1.0)
    sample = np.sort(np.random.default_rng(1).choice(100_000, 1000, replace=False))
    ref = oracle.bh_direct(pm, sample, G=1.0, eps=EPS)
    h = H.barnes_hut(pm, theta, eps=EPS, sample=sample)
    assert _bh_same_as_host(acc[sample], h)
    assert len(bound_violations(acc[sample], h, ref)) == 0
    err = _rel_err(acc[sample].astype(np.float64), ref)
    print(f"theta {theta}: median {np.median(err):.2e} p99 {np.quantile(err, 0.99):.2e} max {err.max():.2e}")


def test_bh_reference_law(ctx):
    """GPRT_BH_REFERENCE_LAW: the reference shader's law.  At theta = 0 it is
    the FP64 direct sum of that law (close pairs and a coincident pair
    included), and it differs from the Plummer result where pairs are close;
    at theta = 0.5 it equals the host build bit for bit."""
    pm = D.plummer(20_000, seed=9)
    pm[1, :3] = pm[0, :3] + np.float32([1e-3, 0, 0])
    pm[2, :3] = pm[0, :3] + np.float32([0, 2e-4, 0])
    pm[4, :3] = pm[3, :3]
    sample = np.concatenate([np.arange(5), np.arange(5, 20_000, 97)])
    acc = ctx.barnes_hut(pm, theta=0.0, eps=EPS, G=1.0, law="reference")
    ref = oracle.bh_direct(pm, sample, G=1.0, eps=EPS, law="reference")
    h = H.barnes_hut(pm, 0.0, eps=EPS, law="reference", sample=sample)
    assert _bh_same_as_host(acc[sample], h)
    err = np.abs(acc[sample].astype(np.float64) - ref).max(axis=1)
    assert np.all(err <= (ROUND_U * U + F64_SLACK) * h["absum"])
    pl = oracle.bh_direct(pm, np.arange(3), G=1.0, eps=EPS)
    assert np.min(_rel_err(acc[:3].astype(np.float64), pl)) > 0.05
    acc5 = ctx.barnes_hut(pm, theta=0.5, eps=EPS, G=1.0, law="reference")
    h5 = H.barnes_hut(pm, 0.5, eps=EPS, law="reference", sample=sample)
    assert _bh_same_as_host(acc5[sample], h5)
    assert len(rounding_violations(acc5[sample], h5)) == 0
    with pytest.raises(ValueError):
        ctx.barnes_hut(pm, law="newton")


def test_bh_two_bodies_and_rejects_negative_mass(ctx):
    pm = np.array([[0, 0, 0, 2.0], [3, 4, 0, 5.0]], np.float32)
    acc = ctx.barnes_hut(pm, theta=0.5, eps=0.5, G=1.5)
    ref = oracle.bh_direct(pm, G=1.5, eps=0.5)
    assert np.allclose(acc, ref, rtol=1e-6)
    bad = pm.copy()
    bad[1, 3] = -1
    with pytest.raises(gprt_hip.GprtError):
        ctx.barnes_hut(bad)


# ---------------------------------------------------------------- the reference FFI
def _legacy_pipe(program):
    L = gprt_hip.lib()
    ir = program.encode()
    p = L.gprt_pipeline_create(ir, len(ir), 0)
    assert p, L.gprt_hip_last_error()
    return L, p


@pytest.mark.parametrize("k,r", [(5, 0.02), (64, 0.05), (65, 0.05), (300, 0.06), (1000, 0.08)])
def test_legacy_register_heap_knn(k, r):
    """The register-heap program: the k nearest with d64 < r^2, ids then float
    distances, padding, lengths; k > 64 runs one query per wavefront."""
    L, p = _legacy_pipe("gprt-hip:knn-heap")
    try:
        pts, q = D.uniform_cloud(20_000, 2_000 if k <= 64 else 400, seed=12)
        r = np.float32(r)
        geom = np.concatenate([pts, np.full((len(pts), 1), r, np.float32)], 1)
        aabbs = np.concatenate([pts - r, pts + r], 1).astype(np.float32)
        bvh = L.gprt_bvh_build(p, geom.ctypes.data, aabbs.ctypes.data, len(pts), geom.nbytes)
        assert bvh
        L.gprt_register_array(p, b"out", 2 * k, len(q))
        qq = np.concatenate([q, np.full((len(q), 1), r, np.float32)], 1)
        L.gprt_execute(p, bvh, qq.ctypes.data, len(q))
        lens = np.empty(len(q), np.uint32)
        flat = np.empty(len(q) * 2 * k, np.uint32)
        L.gprt_retrieve_array_lengths(p, b"out", lens.ctypes.data, len(q))
        L.gprt_retrieve_array_flat(p, b"out", flat.ctypes.data, flat.size)
        flat = flat.reshape(len(q), 2 * k)
        ids, d2 = oracle.knn_brute(pts, q, k, with_d2=True)
        r2 = float(r) * float(r)
        for i in range(len(q)):
            m = int((d2[i] < r2).sum())
            assert lens[i] == m
            assert np.array_equal(flat[i, :m], ids[i, :m])
            assert np.all(flat[i, m:k] == 0xFFFFFFFF)
            assert np.array_equal(flat[i, k:k + m].view(np.float32), d2[i, :m].astype(np.float32))
        L.gprt_bvh_destroy(bvh)
    finally:
        L.gprt_pipeline_destroy(p)


def test_legacy_push_program():
    L, p = _legacy_pipe("gprt-hip:push")
    try:
        pts, q = D.uniform_cloud(20_000, 1_000, seed=13)
        r = np.float32(0.03)
        geom = np.concatenate([pts, np.full((len(pts), 1), r, np.float32)], 1)
        bvh = L.gprt_bvh_build(p, geom.ctypes.data, None, len(pts), geom.nbytes)
        cap = 16
        L.gprt_register_array(p, b"hits", cap, len(q))
        qq = np.concatenate([q, np.full((len(q), 1), r, np.float32)], 1)
        L.gprt_execute(p, bvh, qq.ctypes.data, len(q))
        lens = np.empty(len(q), np.uint32)
        flat = np.empty(len(q) * cap, np.uint32)
        L.gprt_retrieve_array_lengths(p, b"hits", lens.ctypes.data, len(q))
        L.gprt_retrieve_array_flat(p, b"hits", flat.ctypes.data, flat.size)
        off, ids = oracle.range_brute(pts, q, float(r))
        assert np.array_equal(lens.astype(np.int64), np.diff(off))
        flat = flat.reshape(len(q), cap)
        for i in range(len(q)):
            m = min(cap, int(off[i + 1] - off[i]))
            assert np.array_equal(flat[i, :m], ids[off[i]:off[i] + m])
        L.gprt_bvh_destroy(bvh)
    finally:
        L.gprt_pipeline_destroy(p)


def test_legacy_autorope_single_leaf_is_direct_sum():
    L, p = _legacy_pipe("gprt-hip:autorope")
    try:
        rng = np.random.default_rng(3)
        nb = 500
        body = np.zeros(nb, dtype=[("pos", "<f4", 3), ("mass", "<f4"), ("vel", "<f4", 3)])
        body["pos"] = rng.normal(size=(nb, 3))
        body["mass"] = rng.uniform(0.5, 1.5, nb)
        spatial = np.zeros(1, dtype=[("com", "<f4", 3), ("mass", "<f4")])
        routing = np.zeros(1, dtype=[("next", "<u4"), ("rope", "<u4"), ("leaf", "<u4"), ("width", "<f4"),
                                     ("start", "<u4"), ("num", "<u4"), ("pad", "<u4", 2)])
        routing[0] = (1, 1, 1, 10.0, 0, nb, (0, 0))
        out = np.zeros((nb, 3), np.float32)
        L.gprt_execute_autorope_soa_ffi(p, body.ctypes.data, nb, spatial.ctypes.data, 1, routing.ctypes.data, 0.5,
                                        out.ctypes.data)
        x = body["pos"].astype(np.float64)
        m = body["mass"].astype(np.float64)
        diff = x[None] - x[:, None]
        d2 = (diff ** 2).sum(-1)
        w = np.where(d2 > 1e-10, 6.674e-11 * m[:, None] * m[None] / (d2 * np.sqrt(d2 + 1e-6) + (d2 <= 1e-10)), 0)
        ref = (w[..., None] * diff).sum(1)
        assert np.allclose(out, ref, rtol=1e-3, atol=1e-16)
    finally:
        L.gprt_pipeline_destroy(p)


# ---------------------------------------------------------------- autotuner
def test_autotune_writes_schedule(ctx, tmp_path):
    pts, q = D.uniform_cloud(60_000, 20_000, seed=14)
    path = str(tmp_path / "gprt_wisdom.json")
    cfg, leaf = ctx.autotune(pts, q, 5, budget_trials=12, json_path=path)
    assert leaf in (4, 8, 16)
    js = json.load(open(path))
    assert "radius_increment_mult" in js and "radius_heuristic" in js and js["gprt_hip"]["leaf_size"] == leaf
    t = ctx.build(pts, leaf_size=leaf)
    assert np.array_equal(t.k_nn(q, 5, cfg), oracle.knn_brute(pts, q, 5))
    # the written file reads back to the same schedule
    cfg2, leaf2 = gprt_hip.load_schedule(path)
    assert leaf2 == leaf
    for f, _ in gprt_hip.Config._fields_:
        assert getattr(cfg2, f) == pytest.approx(getattr(cfg, f), rel=1e-5), f
    assert js["build_strategy"] == "PreferFastBuild"  # the LBVH-only form


def test_autotune_chooses_tree(ctx, tmp_path):
    """The autotuner with the tree knob: its pick (LBVH or SAH build) gives the
    oracle's ids, and the file's build_strategy reads back to the same tree."""
    rng = np.random.default_rng(21)
    pts = (rng.random((60_000, 3)) * [50.0, 50.0, 0.0]).astype(np.float32)  # a noisy plane
    pts[:, 2] = rng.normal(size=60_000).astype(np.float32) * 0.02
    q = pts[rng.integers(0, len(pts), 20_000)] + rng.normal(size=(20_000, 3)).astype(np.float32) * 0.01
    path = str(tmp_path / "gprt_wisdom.json")
    cfg, leaf, sah = ctx.autotune(pts, q, 5, budget_trials=10, json_path=path, with_tree=True)
    js = json.load(open(path))
    assert js["build_strategy"] == ("PreferFastTrace" if sah else "PreferFastBuild")
    assert gprt_hip.load_schedule(path, with_tree=True)[1:] == (leaf, sah)
    t = ctx.build(pts, leaf_size=leaf, sah=sah)
    try:
        assert np.array_equal(t.k_nn(q, 5, cfg), oracle.knn_brute(pts, q, 5))
    finally:
        t.close()


def test_autotune_batched_uneven_clouds(ctx, tmp_path):
    """The autotuner over an uneven batch (clouds of different sizes, k in
    {1, 5, 16, 64}): its pick, on the tree it chose, gives every cloud the
    oracle's ids, and the written file reads back to the same schedule."""
    rng = np.random.default_rng(31)
    sizes = [300, 5_000, 1_200, 20_000, 64, 9_000]
    ks = [1, 5, 16, 64, 5, 16]
    clouds = [rng.normal(size=(n, 3)).astype(np.float32) * (1 + i) for i, n in enumerate(sizes)]
    qs = [c[rng.integers(0, len(c), max(1, len(c) // 10))] + rng.normal(size=(max(1, len(c) // 10), 3)).astype(
        np.float32) * 0.01 for c in clouds]
    pts = np.concatenate(clouds)
    q = np.concatenate(qs)
    poff = np.r_[0, np.cumsum(sizes)].astype(np.int64)
    qoff = np.r_[0, np.cumsum([len(x) for x in qs])].astype(np.int64)
    kc = np.array(ks, np.int32)
    path = str(tmp_path / "gprt_wisdom.json")
    cfg, leaf, sah = ctx.autotune_batched(pts, poff, q, qoff, kc, budget_trials=10, json_path=path,
                                          with_tree=True)
    assert leaf in (4, 8, 16)
    assert gprt_hip.load_schedule(path, with_tree=True)[1:] == (leaf, sah)
    t = ctx.build_batched(pts, poff, leaf_size=leaf, sah=sah)
    try:
        ids, out_off = t.k_nn_batched(q, qoff, kc, cfg)
    finally:
        t.close()
    for c in range(len(sizes)):
        want = oracle.knn_brute(clouds[c], qs[c], ks[c])
        assert np.array_equal(ids[out_off[c]:out_off[c + 1]].reshape(-1, ks[c]), want), c
