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
