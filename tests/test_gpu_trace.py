"""The legacy generic trace API (gprt_build_index! / gprt_trace! /
gprt_trace_batched!, gprt-macros/src/lib.rs:221-237, 625-727) mirrored in
gprt_hip.trace over the drop-in FFI, against the FP64 oracle's range lists."""
import numpy as np
import pytest

import gprt_hip
import oracle
from gprt_hip import trace as T

pytestmark = pytest.mark.gpu


def _rays(q, r):
    r = np.broadcast_to(np.asarray(r, np.float32), (len(q),))
    return np.concatenate([q, r[:, None]], 1).astype(np.float32)


def _oracle_lists(pts, rays):
    """Per-ray id lists in (d64, id) order, one oracle call per distinct radius."""
    out = [None] * len(rays)
    for r in np.unique(rays[:, 3]):
        sel = np.nonzero(rays[:, 3] == r)[0]
        off, ids = oracle.range_brute(pts, rays[sel, :3], float(r))
        for j, i in enumerate(sel):
            out[i] = ids[off[j]:off[j + 1]]
    return out


def test_trace_batched_mixed_radii():
    rng = np.random.default_rng(31)
    pts = rng.random((30_000, 3), dtype=np.float32)
    q = rng.random((3_000, 3), dtype=np.float32)
    rays = _rays(q, rng.choice(np.array([0.02, 0.05, 0.08], np.float32), len(q)))
    idx = T.build_index(pts, 0.08)
    try:
        cap = 40
        flat, lens, c = T.trace_batched(idx, rays, cap)
        assert c == cap and flat.shape == (len(q) * cap,) and lens.shape == (len(q),)
        want = _oracle_lists(pts, rays)
        flat = flat.reshape(len(q), cap)
        assert any(len(w) > cap for w in want)  # some rays overflow their slots
        for i, w in enumerate(want):
            assert lens[i] == len(w)
            m = min(cap, len(w))
            assert np.array_equal(flat[i, :m], w[:m])
            assert (flat[i, m:] == 0).all()
    finally:
        idx.close()


def test_trace_flat_list_and_truncation():
    rng = np.random.default_rng(32)
    pts = rng.random((20_000, 3), dtype=np.float32)
    q = rng.random((400, 3), dtype=np.float32)
    idx = T.build_index(pts, lambda p: np.concatenate([p, np.full((len(p), 1), 0.04, np.float32)], 1))
    try:
        rays = _rays(q, 0.04)
        got = T.trace(idx, rays)
        want = np.concatenate(_oracle_lists(pts, rays))
        assert len(want) < T.FLAT_CAP
        assert np.array_equal(got, want)
        # more than 1,000,000 hits in all: the list keeps the first 1,000,000
        q2 = rng.random((2_500, 3), dtype=np.float32)
        rays2 = _rays(q2, 0.3)
        got2 = T.trace(idx, rays2)
        want2 = np.concatenate(_oracle_lists(pts, rays2))
        assert len(want2) > T.FLAT_CAP and len(got2) == T.FLAT_CAP
        assert np.array_equal(got2, want2[:T.FLAT_CAP])
    finally:
        idx.close()


def test_trace_empty_and_rebuild():
    rng = np.random.default_rng(33)
    pts = rng.random((5_000, 3), dtype=np.float32)
    idx = T.build_index(pts, 0.02)
    try:
        assert T.trace(idx, np.zeros((0, 4), np.float32)).shape == (0,)
        flat, lens, cap = T.trace_batched(idx, np.zeros((0, 4), np.float32), 8)
        assert flat.size == 0 and lens.size == 0 and cap == 8
        rays = _rays(rng.random((100, 3), dtype=np.float32), 0.02)
        a = T.trace(idx, rays)
        idx.mark_dirty()  # the next trace rebuilds the BVH
        assert np.array_equal(T.trace(idx, rays), a)
        # a ray with radius 0 hits only an identical point
        hit = T.trace(idx, _rays(pts[7:8], 0.0))
        assert 7 in hit.tolist()
    finally:
        idx.close()
    empty = T.build_index(np.zeros((0, 3), np.float32), 0.1)
    try:
        flat, lens, _ = T.trace_batched(empty, _rays(pts[:10], 0.5), 4)
        assert (lens == 0).all() and (flat == 0).all()
    finally:
        empty.close()
