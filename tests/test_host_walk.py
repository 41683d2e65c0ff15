"""The per-lane FP32 k-NN walk of the batched form (lane_walk.hpp: walk,
candidate list, guard band, FP64 re-rank) and the LBVH layout, built for the
host (libgprt_host.so) with every point, node, LDS-ring and spill access
bounds-checked, against the FP64 oracle on the CPU.

The recorded inputs on which the GPU build of the 64- and 67-slot lists
returned wrong ids (integer lattices with exact distance ties, k = 55..64) are
run here through the same source code at those capacities."""
import numpy as np
import pytest

import oracle
from gprt_hip import hostwalk as H
from test_gpu_parity_matrix import _cloud, _queries


def _check(pts, q, k, cap, **kw):
    ids, st, stats = H.lane_knn(pts, q, k, cap, **kw)
    ref = oracle.knn_brute(pts, q, k)
    w = st == 0
    bad = np.nonzero((ids[w] != ref[w]).any(axis=1))[0]
    assert len(bad) == 0, f"k={k} cap={cap} {kw}: {len(bad)} written rows differ"
    assert not (st == 2).any(), "id guard tripped"
    return st, stats


def _recorded(case):
    if case == "n1437":
        rng = np.random.default_rng(90063)
        rng.integers(6), rng.uniform(0, np.log(100_000)), rng.integers(1, 65), rng.choice(6), rng.integers(10)
        pts = _cloud("lattice", 1437, rng)
        return pts, _queries(pts, 500, rng), (55,), dict(leaf_size=32, mult=1.1, rounds=4)
    rng = np.random.default_rng(1064)
    pts = _cloud("lattice", 5000, rng)
    return pts, _queries(pts, 1500, rng), (55, 60, 61, 62, 63, 64), dict(leaf_size=8, mult=3.0, rounds=3)


@pytest.mark.parametrize("case", ["n1437", "n5000"])
def test_recorded_lattice_inputs(case):
    """Every written row equals the oracle and no access leaves its buffer, at
    the 64- and 67-slot capacities the GPU build got wrong; the initial radius
    is the LCG sampler's, as on the GPU."""
    pts, q, ks, kw = _recorded(case)
    r0 = oracle.sampler_lcg(pts, q)
    for k in ks:
        for cap in (64, 67):
            if k + 3 <= cap:
                st, _ = _check(pts, q, k, cap, radius0=r0, **kw)
                assert (st == 0).sum() > len(q) // 2


KINDS = ["uniform", "gauss", "clusters", "plane", "lattice", "line"]


@pytest.mark.parametrize("cap", [8, 16, 24, 32, 40, 48])
def test_host_walk_matrix(cap):
    """Every distribution at every per-lane capacity, k at both ends of the
    class, bounded rounds from a tiny radius, and a stack small enough to
    overflow (those queries go to the exact kernel, status 1)."""
    kmax = cap - 3
    for i, kind in enumerate(KINDS):
        rng = np.random.default_rng(40 + cap + i)
        pts = _cloud(kind, 3000, rng)
        q = _queries(pts, 300, rng)
        for k in sorted({max(1, kmax - 7), kmax}):
            _check(pts, q, k, cap, leaf_size=(4, 8, 16)[i % 3])
            _check(pts, q, k, cap, radius0=1e-4, mult=2.0, rounds=12)
            st, _ = _check(pts, q, k, cap, stack_depth=9)
            assert (st == 0).any()


def test_result_independent_of_radius_and_rounds():
    rng = np.random.default_rng(5)
    pts = _cloud("clusters", 4000, rng)
    q = _queries(pts, 400, rng)
    a, sa, _ = H.lane_knn(pts, q, 13, 16)
    for kw in (dict(radius0=1e-5, mult=1.5, rounds=16), dict(radius0=1e3, rounds=1), dict(radius0=0.3, rounds=2)):
        b, sb, _ = H.lane_knn(pts, q, 13, 16, **kw)
        both = (sa == 0) & (sb == 0)
        assert both.sum() > 300 and np.array_equal(a[both], b[both])


def test_capacity_too_small_is_rejected():
    pts = np.random.default_rng(0).random((100, 3), dtype=np.float32)
    with pytest.raises(RuntimeError):
        H.lane_knn(pts, pts[:4], 6, 8)


@pytest.mark.parametrize("kind,n,leaf", [("uniform", 5000, 8), ("lattice", 1000, 1), ("plane", 777, 4),
                                         ("clusters", 3, 8), ("line", 2, 1)])
def test_host_tree_layout(kind, n, leaf):
    """Keys are the 21-bit-per-axis Morton codes over the cube key box, stably
    sorted; every child box is the min/max of its point range and the child
    counts partition the node's range (a walk of the tree reaches every point
    once).  An internal child's first sorted point follows from its reference
    alone, as the range walk's whole-subtree runs assume: ref - count + 1 for
    an A child (a Karras node ending at the split), ref for a B child."""
    rng = np.random.default_rng(n)
    pts = _cloud(kind, n, rng)
    keys, perm, nodes, root = H.build_tree(pts, leaf)
    lo, hi = pts.min(0), pts.max(0)
    e = np.float32((hi - lo).max())
    kb = np.concatenate([lo, lo + e]).astype(np.float32)
    raw = oracle.morton_codes_wide(pts, kb, 21)
    assert np.array_equal(perm, np.argsort(raw, kind="stable"))
    assert np.array_equal(keys, raw[perm])
    sp = pts[perm]
    seen = np.zeros(n, np.int64)
    stack = [(int(root[0]), int(root[1]), 0)]
    while stack:
        ref, cnt, first = stack.pop()
        if ref < 0:
            assert ~ref == first
            seen[first:first + cnt] += 1
            continue
        nd = nodes[ref]
        ca, cb = int(nd[7:8].view(np.int32)[0]), int(nd[15:16].view(np.int32)[0])
        ra, rb = int(nd[3:4].view(np.int32)[0]), int(nd[11:12].view(np.int32)[0])
        assert ca + cb == cnt
        for (r, c, f, box, side) in ((ra, ca, first, nd[0:8], 0), (rb, cb, first + ca, nd[8:16], 1)):
            if r >= 0:
                assert (r if side else r - c + 1) == f
            seg = sp[f:f + c]
            assert np.array_equal(box[0:3], seg.min(0)) and np.array_equal(box[4:7], seg.max(0))
            stack.append((r, c, f))
    assert (seen == 1).all()
