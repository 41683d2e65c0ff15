"""Device allocation failures: with the library's fault hook
(gprt_hip_debug_fail_alloc, also GPRT_HIP_FAULT=alloc[:N]) the N-th device
allocation and every later one fail.  For every N up to the first that lets
the call through, each entry point must return GPRT_ERR_ALLOC (never crash,
never exit as the reference does on any failure, main.cpp:12-28), the context
must stay usable, and destroying it must free every buffer the library
allocated."""
import ctypes

import numpy as np
import pytest

import gprt_hip
import oracle

pytestmark = pytest.mark.gpu


def _live():
    c, b, n = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    gprt_hip.lib().gprt_hip_debug_alloc_stats(ctypes.byref(c), ctypes.byref(b), ctypes.byref(n))
    return b.value, n.value


RNG = np.random.default_rng(77)
PTS = RNG.random((30_000, 3), dtype=np.float32)
Q = RNG.random((3_000, 3), dtype=np.float32)
POFF = np.array([0, 10_000, 30_000], np.int64)
QOFF = np.array([0, 1_000, 3_000], np.int64)


def _knn(ctx):
    return ctx.build(PTS).k_nn(Q, 5)


def _knn_nograph(ctx):
    return ctx.build(PTS).k_nn(Q, 7, gprt_hip.default_config(use_graph=0, compact_rounds=1))


def _plain(ctx):
    return ctx.build(PTS).k_nn_plain(Q, 5)


def _large_k(ctx):
    return ctx.build(PTS).k_nn(Q[:200], 100)


def _batched(ctx):
    return ctx.build_batched(PTS, POFF).k_nn_batched(Q, QOFF, np.array([3, 50], np.int32))[0]


def _range(ctx):
    return ctx.build(PTS).range_query(Q, 0.03)[1]


def _bh(ctx):
    pm = np.concatenate([PTS, np.full((len(PTS), 1), 1.0 / len(PTS), np.float32)], 1)
    return ctx.barnes_hut(pm, theta=0.5)


def _radius(ctx):
    return np.float32(ctx.sample_radius(PTS, Q))


def _sah(ctx):
    return ctx.build(PTS, sah=True).k_nn(Q, 5)


def _sah_batched(ctx):
    return ctx.build_batched(PTS, POFF, sah=True).k_nn_batched(Q, QOFF, np.array([3, 50], np.int32))[0]


OPS = [_knn, _knn_nograph, _plain, _large_k, _batched, _range, _bh, _radius, _sah, _sah_batched]


@pytest.mark.parametrize("op", OPS, ids=[f.__name__ for f in OPS])
def test_alloc_failures_are_clean(op):
    L = gprt_hip.lib()
    base = _live()
    want = None
    for n in range(200):
        print(f"{op.__name__}: allocations allowed {n}", flush=True)
        L.gprt_hip_debug_fail_alloc(n)
        ctx = None
        try:
            ctx = gprt_hip.Context(0)
            got = op(ctx)
        except gprt_hip.GprtError as e:
            assert e.code == gprt_hip.ERR_ALLOC, str(e)
            got = None
        finally:
            L.gprt_hip_debug_fail_alloc(-1)
        if ctx is not None:
            # the context still works once allocations succeed again
            want = op(ctx)
            if got is not None:
                assert np.array_equal(np.asarray(got), np.asarray(want))
            ctx.close()
        assert _live() == base, f"{op.__name__}: buffers leaked after failing at allocation {n}"
        if got is not None:
            break
    else:
        pytest.fail("the call never succeeded")
    assert n > 0 and want is not None, "no allocation failure was exercised"
