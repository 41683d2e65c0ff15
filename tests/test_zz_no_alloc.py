"""gprt_hip_knn_noalloc, the allocation-free variant of gprt_hip_knn: every
per-call workspace in the library's scratch array, no device allocation, same
ids and distances as gprt_hip_knn (bit for bit) and the FP64 oracle, on two
calls in a row with the second smaller."""
import ctypes

import numpy as np
import pytest

import gprt_hip
import oracle

pytestmark = pytest.mark.gpu


def _stats():
    c, b, n = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    gprt_hip.lib().gprt_hip_debug_alloc_stats(ctypes.byref(c), ctypes.byref(b), ctypes.byref(n))
    return c.value, b.value, n.value


@pytest.mark.parametrize("k", [5, 16, 64])
def test_noalloc_equals_knn(ctx, k):
    L = gprt_hip.lib()
    rng = np.random.default_rng(300 + k)
    pts = rng.random((120_000, 3), dtype=np.float32)
    q1 = (pts[rng.integers(0, len(pts), 60_000)] + rng.normal(size=(60_000, 3)).astype(np.float32) * 1e-3)
    q2 = rng.random((7_000, 3), dtype=np.float32)
    tree = ctx.build(pts, leaf_size=8)
    cfg = gprt_hip.default_config()
    try:
        for q in (q1.astype(np.float32), q2):
            nq = len(q)
            dq = gprt_hip.DeviceArray.from_host(ctx, q)
            ids = gprt_hip.DeviceArray(ctx, (nq, k), np.uint32)
            d2 = gprt_hip.DeviceArray(ctx, (nq, k), np.float64)
            print(f"noalloc k={k} nq={nq}", flush=True)
            before = _stats()
            rc = L.gprt_hip_knn_noalloc(ctx.h, tree.h, dq.ptr, nq, k, ctypes.byref(cfg), ids.ptr, d2.ptr)
            assert rc == 0, L.gprt_hip_last_error()
            assert _stats() == before, "the allocation-free call allocated device memory"
            got, got_d2 = ids.to_host(), d2.to_host()
            want, want_d2 = tree.k_nn(q, k, cfg, with_d2=True)
            assert np.array_equal(got, want)
            assert np.array_equal(got_d2.view(np.uint64), want_d2.view(np.uint64))
            assert np.array_equal(got, oracle.knn_brute(pts, q, k))
            for a in (dq, ids, d2):
                a.free()
    finally:
        tree.close()


def test_noalloc_rejects_too_many_queries(ctx):
    L = gprt_hip.lib()
    m = L.gprt_hip_knn_noalloc_max_queries()
    assert m >= 400_000
    pts = np.random.default_rng(0).random((1000, 3), dtype=np.float32)
    tree = ctx.build(pts)
    dq = gprt_hip.DeviceArray.from_host(ctx, pts)
    ids = gprt_hip.DeviceArray(ctx, (len(pts), 5), np.uint32)
    try:
        rc = L.gprt_hip_knn_noalloc(ctx.h, tree.h, dq.ptr, m + 1, 5, None, ids.ptr, None)
        assert rc == gprt_hip.ERR_CAPACITY and b"at most" in L.gprt_hip_last_error()
        # still usable afterwards, and the ordinary entry point too
        assert L.gprt_hip_knn_noalloc(ctx.h, tree.h, dq.ptr, len(pts), 5, None, ids.ptr, None) == 0
        assert np.array_equal(ids.to_host(), oracle.knn_brute(pts, pts, 5))
        assert np.array_equal(tree.k_nn(pts, 5), oracle.knn_brute(pts, pts, 5))
    finally:
        dq.free()
        ids.free()
        tree.close()
