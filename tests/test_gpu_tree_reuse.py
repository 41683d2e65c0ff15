"""A destroyed tree's device buffers go back to its context and the next build
takes them over.  Rebuilding at other sizes, with the other key width, and
with a second tree alive must give the oracle's k-NN ids every time, with the
captured graph (a new tree always forces a new capture, even when its buffers
sit at the old addresses)."""
import numpy as np
import pytest

import gprt_hip
import oracle

pytestmark = pytest.mark.gpu


def test_rebuilds_reuse_buffers(ctx):
    rng = np.random.default_rng(77)
    q = rng.random((3_000, 3), dtype=np.float32)
    cfg = gprt_hip.default_config(use_graph=1)
    keep = ctx.build(rng.random((5_000, 3), dtype=np.float32))
    keep_pts = None
    for step, (n, ref30) in enumerate([(40_000, False), (40_000, False), (9_000, False), (60_000, False),
                                       (60_000, True), (60_000, False), (1, False), (25_000, True)]):
        pts = (rng.random((n, 3), dtype=np.float32) * (1 + step)).astype(np.float32)
        t = ctx.build(pts, ref_keys=ref30)
        want = oracle.knn_brute(pts, q * np.float32(1 + step), 5 if n >= 5 else 1)
        k = want.shape[1]
        for _ in range(2):  # the second call replays the graph captured by the first
            got = t.k_nn(q * np.float32(1 + step), k, cfg)
            assert np.array_equal(got, want), f"step {step}: n={n} ref30={ref30}"
        t.close()
        if step == 3:
            keep_pts = rng.random((7_000, 3), dtype=np.float32)
            keep.close()
            keep = ctx.build(keep_pts)
    assert np.array_equal(keep.k_nn(q, 4, cfg), oracle.knn_brute(keep_pts, q, 4))
    keep.close()


def test_tree_outlives_its_context():
    """A context destroyed before its trees: they keep working through another
    context and free their own buffers (no use of the destroyed context)."""
    rng = np.random.default_rng(78)
    pts = rng.random((20_000, 3), dtype=np.float32)
    q = rng.random((1_000, 3), dtype=np.float32)
    c1 = gprt_hip.Context(0)
    t = c1.build(pts)
    dead = c1.build(rng.random((3_000, 3), dtype=np.float32))
    c1.close()
    c2 = gprt_hip.Context(0)
    try:
        t.ctx = c2
        assert np.array_equal(t.k_nn(q, 6), oracle.knn_brute(pts, q, 6))
        with pytest.raises(gprt_hip.GprtError):
            t.export()
        dead.close()
        t.close()
        t2 = c2.build(pts)
        assert np.array_equal(t2.k_nn(q, 6), oracle.knn_brute(pts, q, 6))
        t2.close()
    finally:
        c2.close()


@pytest.mark.parametrize("kind,n,leaf", [("uniform", 50_000, 8), ("lattice", 5000, 4), ("plane", 3000, 16),
                                         ("clusters", 20_000, 1), ("line", 7, 8)])
def test_gpu_tree_equals_host_build(ctx, kind, n, leaf):
    """The GPU LBVH (keys, permutation, every node's child boxes, refs and
    counts, root) equals the host build of the same layout bit for bit."""
    from gprt_hip import hostwalk as H
    from test_gpu_parity_matrix import _cloud
    pts = _cloud(kind, n, np.random.default_rng(n))
    t = ctx.build(pts, leaf_size=leaf)
    keys, perm, nodes, roots = t.export()
    hk, hp, hn, hr = H.build_tree(pts, leaf)
    t.close()
    assert np.array_equal(keys, hk) and np.array_equal(perm, hp)
    assert np.array_equal(nodes.view(np.uint32), hn.view(np.uint32))
    assert np.array_equal(roots[0], hr)
