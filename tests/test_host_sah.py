"""The SAH tree's host restatement (host_walk.cpp build_tree_sah, the split
decisions of sah.hpp): structural invariants that hold for any correct SAH tree
in the LBVH node layout, and the per-lane walk over it against the FP64 oracle.

The invariants are not a restatement of the build: every point in exactly one
leaf bucket, each reachable internal node's range the concatenation of its
children's, boxes equal to the min/max of their points, node ids in the Karras
numbering (root 0, an internal A child its last point, a B child its first),
parent links and ranges consistent with the references, and a split
separating its children along one axis."""
import numpy as np
import pytest

import oracle
from gprt_hip import hostwalk as H
from test_gpu_parity_matrix import _cloud


def _check_tree(pts, leaf):
    n = len(pts)
    keys, perm, nodes, root, parent, ranges = H.build_tree_ex(pts, leaf, sah=True)
    assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.uint32))
    sp = pts[perm]  # the points in tree order
    if n <= leaf:
        assert root[0] == ~0 and root[1] == n
    else:
        assert root[0] == 0 and root[1] == n
    covered = np.zeros(n, np.int32)
    stack = [(int(root[0]), int(root[1]), 0)]
    internal = 0
    while stack:
        ref, cnt, first = stack.pop()
        if ref < 0:
            assert ~ref == first
            assert cnt <= leaf or n <= leaf
            covered[first:first + cnt] += 1
            continue
        internal += 1
        nd = nodes[ref]
        ca, cb = int(nd[7].view(np.int32)), int(nd[15].view(np.int32))
        assert ca >= 1 and cb >= 1 and ca + cb == cnt
        assert ranges[ref][0] == first and ranges[ref][1] == first + cnt - 1
        for side, (c, f) in enumerate([(ca, first), (cb, first + ca)]):
            lo, hi = nd[8 * side:8 * side + 3], nd[8 * side + 4:8 * side + 7]
            seg = sp[f:f + c]
            assert np.array_equal(lo, seg.min(0)) and np.array_equal(hi, seg.max(0))
            r = int(nd[8 * side + 3].view(np.int32))
            if c <= leaf:
                assert r == ~f
            else:
                # Karras numbering: an internal A child is its last point, B its first
                assert r == (f + c - 1 if side == 0 else f) and parent[r] == (ref << 1 | side)
            stack.append((r, c, f))
        # the two sides are separated along some axis (a plane split), unless
        # the segment is degenerate (all its points equal) and split at its median
        a, b = sp[first:first + ca], sp[first + ca:first + cnt]
        sep = [a[:, k].max() <= b[:, k].min() for k in range(3)]
        assert any(sep) or (sp[first:first + cnt] == sp[first]).all()
    assert (covered == 1).all()
    # every point's parent link names a node whose range holds it
    for i in range(n):
        pl = parent[(n - 1) + i] if n > 1 else -1
        if pl >= 0:
            r = ranges[pl >> 1]
            assert r[0] <= i <= r[1]
    return keys, perm, nodes, root, internal


@pytest.mark.parametrize("kind", ["uniform", "gauss", "clusters", "plane", "lattice", "line"])
@pytest.mark.parametrize("n,leaf", [(1, 8), (2, 1), (7, 8), (9, 8), (100, 1), (1000, 4), (5000, 8), (3000, 32)])
def test_sah_tree_invariants(kind, n, leaf):
    pts = _cloud(kind, n, np.random.default_rng(n + leaf))
    _check_tree(pts, leaf)


def test_sah_tree_duplicates():
    rng = np.random.default_rng(5)
    pts = np.repeat(rng.random((12, 3), dtype=np.float32), 40, axis=0)
    _check_tree(pts, 8)
    _check_tree(np.zeros((300, 3), np.float32), 8)  # one location: median splits only


def test_sah_tree_deterministic():
    pts = _cloud("clusters", 4000, np.random.default_rng(3))
    a = H.build_tree_ex(pts, 8, sah=True)
    b = H.build_tree_ex(pts, 8, sah=True)
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))


def test_sah_tree_fewer_visits_on_a_surface():
    """On points scattered over two noisy planes the SAH tree needs fewer node
    visits per query than the LBVH (the reason for the option)."""
    rng = np.random.default_rng(11)
    n = 20000
    p = rng.random((n, 3)).astype(np.float32) * 20.0
    p[: n // 2, 2] = rng.normal(size=n // 2).astype(np.float32) * 0.02
    p[n // 2:, 0] = 10.0 + rng.normal(size=n - n // 2).astype(np.float32) * 0.02
    q = p[rng.integers(0, n, 2000)] + rng.normal(size=(2000, 3)).astype(np.float32) * 0.01
    _, _, lb = H.lane_knn(p, q, 5, 8, leaf_size=8, radius0=0.5, sah=False)
    _, _, sa = H.lane_knn(p, q, 5, 8, leaf_size=8, radius0=0.5, sah=True)
    assert sa[0] < lb[0]


@pytest.mark.parametrize("kind", ["uniform", "clusters", "plane", "lattice"])
@pytest.mark.parametrize("k,cap", [(1, 8), (5, 8), (16, 24), (32, 40)])
def test_lane_walk_on_sah_tree_exact(kind, k, cap):
    """The per-lane FP32 walk with FP64 re-rank over the SAH tree returns the
    oracle's ids for every query it resolves, and never reads out of bounds."""
    rng = np.random.default_rng(k * 7 + cap)
    pts = _cloud(kind, 6000, rng)
    span = float(np.ptp(pts, axis=0).max())
    q = (pts[rng.integers(0, len(pts), 1500)] + rng.normal(size=(1500, 3)).astype(np.float32) * 0.01 * span)
    q = q.astype(np.float32)
    ref = oracle.knn_brute(pts, q, k)
    ids, st, _ = H.lane_knn(pts, q, k, cap, leaf_size=8, radius0=0.02 * span, sah=True)
    assert not (st == 2).any()
    ok = st == 0
    assert ok.mean() > 0.9
    assert np.array_equal(ids[ok], ref[ok])


def _deep_cloud(n=2000):
    """Geometrically spaced points along x: every SAH split peels off the far
    end (about a hundred points a level), a deep tree for its size."""
    x = np.power(np.float64(1.02), np.arange(n))
    p = np.zeros((n, 3), np.float32)
    p[:, 0] = x.astype(np.float32)
    return p


def test_sah_depth_capped():
    pts = _deep_cloud()
    n = len(pts)
    _check_tree(pts, 8)
    _, _, nodes, root, parent, ranges = H.build_tree_ex(pts, 8, sah=True)
    depth, stack = 0, [(int(root[0]), 0)]
    while stack:
        r, d = stack.pop()
        depth = max(depth, d)
        if r >= 0:
            for side in range(2):
                stack.append((int(nodes[r][8 * side + 3].view(np.int32)), d + 1))
    # SAH splits stop at depth 48, index-median splits below halve the segment
    assert 20 < depth <= 48 + int(np.ceil(np.log2(n))) + 1
