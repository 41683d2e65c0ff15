"""SAH trees (GPRT_BUILD_SAH) on the GPU: the built tree equals the host
restatement bit for bit (sah.hip against host_walk.cpp build_tree_sah, the
same decisions of sah.hpp), and every query type over it returns the oracle's
results (queries do not depend on the tree's shape)."""
import numpy as np
import pytest

import gprt_hip
import oracle
from gprt_hip import datasets as D
from test_gpu_parity_matrix import _cloud, _queries

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,n,leaf", [("uniform", 50_000, 8), ("lattice", 5000, 4), ("plane", 3000, 16),
                                         ("clusters", 20_000, 8), ("line", 4000, 1), ("gauss", 2500, 32),
                                         ("uniform", 9, 8), ("uniform", 2, 1), ("clusters", 1500, 8)])
def test_sah_tree_equals_host_build(ctx, kind, n, leaf):
    """Keys, permutation, every node's child boxes, references and counts, and
    the root equal the host build of the SAH tree bit for bit."""
    from gprt_hip import hostwalk as H
    pts = _cloud(kind, n, np.random.default_rng(n + 17))
    print(f"sah tree {kind} n={n} leaf={leaf}", flush=True)
    t = ctx.build(pts, leaf_size=leaf, sah=True)
    try:
        keys, perm, nodes, roots = t.export()
    finally:
        t.close()
    hk, hp, hn, hr, _, _ = H.build_tree_ex(pts, leaf, sah=True)
    assert np.array_equal(perm, hp)
    assert np.array_equal(keys, hk)
    assert np.array_equal(nodes.view(np.uint32), hn.view(np.uint32))
    assert np.array_equal(roots[0], hr)


def test_sah_tree_duplicates_equal_host(ctx):
    from gprt_hip import hostwalk as H
    rng = np.random.default_rng(5)
    for pts in (np.repeat(rng.random((12, 3), dtype=np.float32), 300, axis=0), np.zeros((3000, 3), np.float32)):
        t = ctx.build(pts, leaf_size=8, sah=True)
        try:
            keys, perm, nodes, roots = t.export()
        finally:
            t.close()
        hk, hp, hn, hr, _, _ = H.build_tree_ex(pts, 8, sah=True)
        assert np.array_equal(perm, hp) and np.array_equal(roots[0], hr)
        assert np.array_equal(nodes.view(np.uint32), hn.view(np.uint32))


@pytest.mark.parametrize("kind", ["uniform", "clusters", "plane", "lattice", "line"])
@pytest.mark.parametrize("k", [1, 5, 16, 32, 64, 100])
def test_sah_knn_parity(ctx, kind, k):
    rng = np.random.default_rng(k * 31 + len(kind))
    pts = _cloud(kind, 20_000, rng)
    q = _queries(pts, 3000, rng)
    ref = oracle.knn_brute(pts, q, k)
    print(f"sah knn {kind} k={k}", flush=True)
    t = ctx.build(pts, leaf_size=8, sah=True)
    try:
        assert np.array_equal(t.k_nn(q, k), ref), "batched form"
        assert np.array_equal(t.k_nn_plain(q, k), ref), "plain form"
        cfg = gprt_hip.default_config(compact_rounds=1, traversal=gprt_hip.TRAV_WIDE)
        assert np.array_equal(t.k_nn(q, k, cfg), ref), "4-wide walk, compacted rounds"
    finally:
        t.close()


@pytest.mark.parametrize("kind,rel", [("uniform", 0.01), ("clusters", 0.05), ("plane", 0.02), ("lattice", 0.2)])
def test_sah_range_parity(ctx, kind, rel):
    rng = np.random.default_rng(77)
    pts = _cloud(kind, 8000, rng)
    q = _queries(pts, 800, rng)
    span = float(np.ptp(pts, axis=0).max())
    r = float(np.float32(rel * span))
    roff, rids = oracle.range_brute(pts, q, r)
    t = ctx.build(pts, leaf_size=8, sah=True)
    try:
        off, ids = t.range_query(q, r)
    finally:
        t.close()
    assert np.array_equal(off, roff) and np.array_equal(ids, rids)


def test_sah_multi_cloud(ctx):
    """Uneven clouds (empty, single points, at most a leaf, thousands of
    points) in one SAH tree: per-cloud k-NN equals each cloud's oracle."""
    rng = np.random.default_rng(9)
    sizes = [0, 1, 5, 8, 9, 2000, 70_000, 3, 1500, 0, 30_000]
    clouds = [_cloud(("uniform", "clusters", "plane")[i % 3], s, rng) for i, s in enumerate(sizes)]
    pts = np.concatenate([c for c in clouds if len(c)] or [np.zeros((0, 3), np.float32)]).astype(np.float32)
    poff = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    qs = [_queries(c, 200, rng) if len(c) else rng.random((20, 3), dtype=np.float32) for c in clouds]
    q = np.concatenate(qs).astype(np.float32)
    qoff = np.concatenate([[0], np.cumsum([len(x) for x in qs])]).astype(np.int64)
    kc = np.array([5, 1, 5, 16, 5, 32, 5, 2, 64, 5, 8], np.int32)
    t = ctx.build_batched(pts, poff, leaf_size=8, sah=True)
    try:
        ids, out_off = t.k_nn_batched(q, qoff, kc, gprt_hip.default_config())
    finally:
        t.close()
    for c in range(len(sizes)):
        k = int(kc[c])
        ref = oracle.knn_brute(pts[poff[c]:poff[c + 1]], q[qoff[c]:qoff[c + 1]], k)
        got = ids[out_off[c]:out_off[c + 1]].reshape(-1, k)
        assert np.array_equal(got, ref), f"cloud {c} (n={sizes[c]}, k={k})"


def test_sah_lidar_full(ctx):
    """The LiDAR workload (1.31 M points, 400 K queries) at k = 5 and 32 on the
    SAH tree: every query's ids equal the oracle's."""
    pts, q = D.lidar_scan()
    t = ctx.build(pts, leaf_size=8, sah=True)
    try:
        for k in (5, 32):
            ref, _, _, _ = oracle.kdtree_knn(pts, q, k)
            got = t.k_nn(q, k)
            bad = np.nonzero((got != ref).any(1))[0]
            assert len(bad) == 0, f"k={k}: {len(bad)} queries differ, first {bad[:5]}"
    finally:
        t.close()


def test_sah_rebuild_reuses_and_matches(ctx):
    """A SAH build after an LBVH build on the same context (reused buffers) and
    the reverse give the same trees as fresh builds."""
    pts = _cloud("clusters", 30_000, np.random.default_rng(2))
    t1 = ctx.build(pts, leaf_size=8, sah=True)
    a = t1.export()
    t1.close()
    t2 = ctx.build(pts, leaf_size=8)
    b = t2.export()
    t2.close()
    t3 = ctx.build(pts, leaf_size=8, sah=True)
    c = t3.export()
    t3.close()
    for x, y in zip(a, c):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    assert not np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32))


def test_sah_deep_tree_equals_host(ctx):
    """Geometrically spaced points (a deep SAH tree): GPU tree equals the host's,
    and k-NN over it equals the oracle."""
    from gprt_hip import hostwalk as H
    from test_host_sah import _deep_cloud
    pts = _deep_cloud()
    t = ctx.build(pts, leaf_size=8, sah=True)
    try:
        keys, perm, nodes, roots = t.export()
        q = pts[::7] * np.float32(1.0001)
        ref = oracle.knn_brute(pts, q, 5)
        assert np.array_equal(t.k_nn(q, 5), ref)
    finally:
        t.close()
    hk, hp, hn, hr, _, _ = H.build_tree_ex(pts, 8, sah=True)
    assert np.array_equal(perm, hp) and np.array_equal(roots[0], hr)
    assert np.array_equal(nodes.view(np.uint32), hn.view(np.uint32))
