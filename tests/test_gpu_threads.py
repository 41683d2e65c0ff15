"""Calls on one context from several host threads.  Each C-ABI call holds the
context's mutex until it returns (ctypes releases the GIL, so the threads do
enter the library together); every thread's results must still equal the FP64
oracle's, in k-NN of several candidate classes, range queries and tree builds
that all share the context's staging buffers."""
import threading

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_threads_share_one_context(ctx):
    rng = np.random.default_rng(41)
    pts = rng.random((20_000, 3), dtype=np.float32)
    pts2 = rng.random((7_000, 3), dtype=np.float32)
    tree = ctx.build(pts)
    qs = [rng.random((3_000 + 500 * i, 3), dtype=np.float32) for i in range(4)]
    ks = [1, 5, 16, 33]
    want = [oracle.knn_brute(pts, q, k) for q, k in zip(qs, ks)]
    rq = qs[0][:500]
    roff, rids = oracle.range_brute(pts, rq, 0.03)
    want2 = oracle.knn_brute(pts2, qs[1][:800], 4)
    errors = []

    def worker(i):
        try:
            for rep in range(6):
                if i == 3 and rep % 2:
                    off, ids = tree.range_query(rq, 0.03)
                    assert np.array_equal(off, roff) and np.array_equal(ids, rids), f"range, rep {rep}"
                elif i == 2 and rep % 3 == 1:
                    t2 = ctx.build(pts2)
                    try:
                        assert np.array_equal(t2.k_nn(qs[1][:800], 4), want2), f"second tree, rep {rep}"
                    finally:
                        t2.close()
                else:
                    got = tree.k_nn(qs[i], ks[i])
                    assert np.array_equal(got, want[i]), f"thread {i} k={ks[i]} rep {rep}"
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    tree.close()
    assert not any(t.is_alive() for t in threads), "a worker thread did not finish"
    assert not errors, errors[0]
