"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each function is checked against something other than its own formula:
closed forms, the committed golden fixtures (scipy cKDTree, Python integer
arithmetic), pure-Python / numpy restatements written independently, brute
force on tiny inputs, and planted mistakes that the checks must catch.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- Morton
def test_morton_closed_forms():
    assert oracle.expand_bits(0x3FF) == 0x09249249
    assert oracle.expand_bits(1) == 1
    assert oracle.expand_bits(2) == 8
    assert oracle.expand_bits(0x400) == 0  # only 10 bits are kept
    with open(os.path.join(GOLD, "morton_known.json")) as f:
        for (x, y, z), code in json.load(f):
            assert oracle.morton3(x, y, z) == code, (x, y, z)


def _py_morton(x, y, z):
    c = 0
    for i in range(10):
        c |= ((x >> i) & 1) << (3 * i + 2) | ((y >> i) & 1) << (3 * i + 1) | ((z >> i) & 1) << (3 * i)
    return c


def _py_codes(pts):
    """Independent numpy-float32 restatement of morton.rs:19-48."""
    p = pts.astype(np.float32)
    mn = p.min(0)
    mx = p.max(0)
    ext = np.maximum(mx - mn, np.float32(1e-6)).astype(np.float32)
    out = []
    for r in p:
        cells = []
        for a in range(3):
            t = np.float32(np.float32(r[a] - mn[a]) / ext[a]) * np.float32(1023.0)
            t = np.float32(t)
            u = 0 if not (t > 0) else int(t)
            cells.append(min(u, 1023))
        out.append(_py_morton(*cells))
    return np.array(out, np.uint32)


@pytest.mark.parametrize("case", ["uniform", "flat_axis", "negative", "big"])
def test_morton_codes_match_python(case):
    rng = np.random.default_rng(5)
    pts = rng.random((500, 3), dtype=np.float32)
    if case == "flat_axis":
        pts[:, 1] = 3.25  # zero extent -> 1e-6
    elif case == "negative":
        pts = pts * 200 - 150
    elif case == "big":
        pts = pts * 1e5
    codes, bbox = oracle.morton_codes(pts)
    assert np.array_equal(codes, _py_codes(pts))
    assert np.array_equal(bbox[:3], pts.min(0)) and np.array_equal(bbox[3:], pts.max(0))
    # the max corner maps to cell 1023 on every axis with extent
    top = np.argmax(pts[:, 0])
    assert (codes[top] >> 2) & 1 == 1


def test_morton_query_outside_box_saturates():
    bbox = np.array([0, 0, 0, 1, 1, 1], np.float32)
    q = np.array([[-5, -5, -5], [5, 5, 5], [0.5, 2.0, -1.0]], np.float32)
    codes = oracle.morton_codes_bbox(q, bbox)
    assert codes[0] == 0
    assert codes[1] == 0x3FFFFFFF
    assert codes[2] == _py_morton(int(np.float32(0.5) * np.float32(1023)), 1023, 0)


def test_morton_wide_keys():
    """The tree's sort keys: at 10 bits per axis they are the reference's
    codes; at other widths each axis de-interleaves to the float32 cell index."""
    rng = np.random.default_rng(8)
    pts = (rng.random((400, 3), dtype=np.float32) * 300 - 100).astype(np.float32)
    pts[0] = [-100, -100, -100]
    pts[1] = [500, 500, 500]  # outside the box: saturates
    bbox = np.array([-100, -100, -100, 200, 200, 200], np.float32)
    assert np.array_equal(oracle.morton_codes_wide(pts, bbox, 10), oracle.morton_codes_bbox(pts, bbox).astype(np.uint64))
    for bits in (1, 7, 16, 21):
        w = oracle.morton_codes_wide(pts, bbox, bits)
        qmax = (1 << bits) - 1
        t = (pts - bbox[:3]) / (bbox[3:] - bbox[:3]) * np.float32(qmax)
        cell = np.minimum(np.where(t > 0, t, 0).astype(np.uint64), qmax)
        for a in range(3):
            got = np.zeros(len(pts), np.uint64)
            for lvl in range(bits):
                got |= ((w >> np.uint64(3 * lvl + 2 - a)) & np.uint64(1)) << np.uint64(lvl)
            assert np.array_equal(got, cell[:, a]), (bits, a)
        assert w[1] == (1 << (3 * bits)) - 1 and w[0] == 0
    with pytest.raises(ValueError):
        oracle.morton_codes_wide(pts, bbox, 22)


# ---------------------------------------------------------------- LCG sampler
def test_lcg_draws_golden():
    with open(os.path.join(GOLD, "lcg_draws.json")) as f:
        g = json.load(f)
    d = oracle.lcg_draws(int(g["seed"], 16), 32)
    assert [int(x) for x in d] == g["draws"]


def _py_sampler_lcg(pts, q, heuristic, p):
    """Pure-Python restatement of dispatch.rs:65-103 (float32 scalars)."""
    f32 = np.float32
    seed = 0x123456789ABCDEF
    n, nq = len(pts), len(q)
    S, QS = min(5000, n), min(100, nq)
    kth = []
    MAX = np.finfo(np.float32).max
    for _ in range(QS):
        seed = (seed * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        qi = (seed >> 33) % nq
        d = [MAX] * 4
        for _ in range(S):
            seed = (seed * 6364136223846793005 + 1442695040888963407) % (1 << 64)
            pi = (seed >> 33) % n
            dx, dy, dz = f32(q[qi, 0] - pts[pi, 0]), f32(q[qi, 1] - pts[pi, 1]), f32(q[qi, 2] - pts[pi, 2])
            dd = np.sqrt(f32(f32(f32(dx * dx) + f32(dy * dy)) + f32(dz * dz)))
            if dd < d[3]:
                d[3] = dd
                if d[3] < d[2]: d[2], d[3] = d[3], d[2]
                if d[2] < d[1]: d[1], d[2] = d[2], d[1]
                if d[1] < d[0]: d[0], d[1] = d[1], d[0]
        if d[3] != MAX:
            kth.append(d[3])
    r = f32(10.0)
    if kth:
        kth.sort()
        if heuristic == oracle.HEUR_MAX:
            r = f32(kth[-1] * f32(1.1))
        elif heuristic == oracle.HEUR_PERCENTILE:
            idx = int(min(f32(f32(len(kth)) * f32(p)), f32(len(kth) - 1)))
            r = f32(kth[idx] * f32(1.1))
        else:
            r = f32(p)
    return float(r)


@pytest.mark.parametrize("heur,p", [(1, 0.95), (1, 0.10), (0, 0.0), (2, 0.75), (1, 1.0)])
def test_sampler_lcg_matches_python(heur, p):
    rng = np.random.default_rng(21)
    pts = rng.random((900, 3), dtype=np.float32)
    q = rng.random((150, 3), dtype=np.float32)
    assert oracle.sampler_lcg(pts, q, heur, p) == _py_sampler_lcg(pts, q, heur, p)


def test_sampler_lcg_tiny_inputs():
    # fewer than 4 points: dists[3] stays f32::MAX, nothing pushed -> 10.0,
    # and Fixed(r) is ignored in that case (dispatch.rs:92-103)
    pts = np.zeros((3, 3), np.float32)
    q = np.ones((5, 3), np.float32)
    assert oracle.sampler_lcg(pts, q, oracle.HEUR_PERCENTILE, 0.5) == 10.0
    assert oracle.sampler_lcg(pts, q, oracle.HEUR_FIXED, 0.5) == 10.0


def _mmix_draws(count):
    """Knuth's MMIX LCG from the sampler's seed (dispatch.rs:65-69), big-int
    arithmetic; its first 32 draws are the golden file's."""
    seed, out = 0x123456789ABCDEF, []
    for _ in range(count):
        seed = (seed * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        out.append(seed >> 33)
    return out


@pytest.mark.parametrize("n,nq", [(5, 500), (7_000, 137), (4, 100)])
def test_sampler_lcg_closed_form(n, nq):
    """Every data point at the origin and query j at (x_j, 0, 0) with x_j = j + 1:
    whichever points a drawn query samples, its four smallest distances are all
    exactly x_j, so the 4th is x_j.  The draw order (one query draw, then
    min(5000, n) point draws, 100 times; query = draw mod nq) fixes which x_j
    are pushed, and the radius follows in closed form: the p-th entry of the
    sorted list (index min(len*p, len-1) in f32) times 1.1, the last one for
    SampledMax, the parameter itself for Fixed (dispatch.rs:65-103)."""
    pts = np.zeros((n, 3), np.float32)
    q = np.zeros((nq, 3), np.float32)
    q[:, 0] = np.arange(1, nq + 1, dtype=np.float32)
    s = min(5000, n)
    draws = _mmix_draws(100 * (s + 1))
    kth = sorted(np.float32(draws[j * (s + 1)] % nq + 1) for j in range(min(100, nq)))
    f32 = np.float32
    for p in (0.95, 0.5, 0.0, 1.0, 0.123):
        idx = int(min(f32(f32(len(kth)) * f32(p)), f32(len(kth) - 1)))
        assert oracle.sampler_lcg(pts, q, oracle.HEUR_PERCENTILE, p) == float(f32(kth[idx] * f32(1.1))), p
    assert oracle.sampler_lcg(pts, q, oracle.HEUR_MAX, 0.0) == float(f32(kth[-1] * f32(1.1)))
    assert oracle.sampler_lcg(pts, q, oracle.HEUR_FIXED, 0.75) == float(f32(0.75))
    # a planted mistake in the interleaving (query draws taken consecutively)
    # gives another list, so the check above can tell the orders apart
    wrong = sorted(np.float32(draws[j] % nq + 1) for j in range(min(100, nq)))
    assert wrong != kth


@pytest.mark.parametrize("k", [1, 2, 3, 5, 8])
def test_sampler_strided_closed_form(k):
    """1,000 points on 100 parallel lines, ten per line with spacing s_i =
    1 + i/4 (lines 1,000 apart), and 100 queries, query i exactly on the
    fifth point of line i.  Every point and every query is sampled (strides 1).
    The query's own point has d^2 = 0 and is dropped by the d^2 > 1e-10 filter,
    its neighbours lie at 1, 1, 2, 2, ... spacings, so its k-th distance is
    ceil(k/2) * s_i exactly, and the radius is the p-th of those times 1.1
    (gprt-macros/src/lib.rs:425-450)."""
    pts = np.zeros((1000, 3), np.float32)
    q = np.zeros((100, 3), np.float32)
    f32 = np.float32
    for i in range(100):
        sp = f32(1 + i / 4)
        pts[10 * i:10 * i + 10, 0] = np.arange(10, dtype=np.float32) * sp
        pts[10 * i:10 * i + 10, 1] = f32(1000 * i)
        q[i] = pts[10 * i + 4]
    kth = sorted(f32((k + 1) // 2) * f32(1 + i / 4) for i in range(100))
    for p in (0.10, 0.5, 0.95, 0.0, 1.0):
        idx = min(int(f32(f32(100) * f32(p))), 99)
        assert oracle.sampler_strided(pts, q, k, oracle.HEUR_PERCENTILE, p) == float(f32(kth[idx] * f32(1.1))), p
    # a planted mistake, the self-filter left out: the query's own point counts
    # as the nearest, the k-th becomes ceil((k-1)/2) spacings, which the closed
    # form tells apart for odd k
    planted = sorted(f32(k // 2) * f32(1 + i / 4) for i in range(100))
    assert (planted != kth) == (k % 2 == 1)
    # the filter keeps a distinct point at d^2 = 1e-8 > 1e-10
    q2 = q.copy()
    q2[:, 2] = f32(1e-4)
    r2 = oracle.sampler_strided(pts, q2, 1, oracle.HEUR_PERCENTILE, 0.5)
    assert abs(r2 - float(f32(1e-4) * f32(1.1))) < 1e-9


def _py_sampler_strided(pts, q, k, heuristic, p):
    """Pure-Python restatement of gprt-macros/src/lib.rs:417-450."""
    f32 = np.float32
    if heuristic == oracle.HEUR_FIXED:
        return float(f32(p))
    if heuristic != oracle.HEUR_PERCENTILE:
        return 1.0
    n, nq = len(pts), len(q)
    qs, ps = max(nq // 100, 1), max(n // 1000, 1)
    MAX = np.finfo(np.float32).max
    dists = []
    for qi in list(range(0, nq, qs))[:100]:
        knn = [MAX] * k
        for pi in list(range(0, n, ps))[:1000]:
            dx, dy, dz = f32(q[qi, 0] - pts[pi, 0]), f32(q[qi, 1] - pts[pi, 1]), f32(q[qi, 2] - pts[pi, 2])
            d2 = f32(f32(f32(dx * dx) + f32(dy * dy)) + f32(dz * dz))
            if d2 > f32(1e-10) and d2 < knn[k - 1]:
                knn[k - 1] = d2
                knn.sort()
        if knn[k - 1] != MAX:
            dists.append(np.sqrt(f32(knn[k - 1])))
    if not dists:
        return 1.0
    dists.sort()
    idx = min(int(f32(f32(len(dists)) * f32(p))), len(dists) - 1)
    return float(f32(dists[idx] * f32(1.1)))


@pytest.mark.parametrize("k,heur,p", [(5, 1, 0.10), (1, 1, 0.5), (16, 1, 0.95), (5, 2, 2.5), (5, 0, 0.3)])
def test_sampler_strided_matches_python(k, heur, p):
    rng = np.random.default_rng(22)
    pts = rng.random((2500, 3), dtype=np.float32)
    pts[7] = pts[3]  # an exact duplicate: d^2 = 0 is excluded by the 1e-10 filter
    q = np.concatenate([pts[:40], rng.random((160, 3), dtype=np.float32)])
    assert oracle.sampler_strided(pts, q, k, heur, p) == _py_sampler_strided(pts, q, k, heur, p)


# ---------------------------------------------------------------- exact k-NN
def test_knn_golden_scipy():
    g = np.load(os.path.join(GOLD, "knn_uniform.npz"))
    ids, d2 = oracle.knn_brute(g["pts"], g["q"], 5, with_d2=True)
    assert np.array_equal(ids, g["ids"])
    assert np.allclose(np.sqrt(d2), g["dist"], rtol=1e-12, atol=0)


def test_knn_golden_lattice_ties():
    g = np.load(os.path.join(GOLD, "knn_lattice.npz"))
    assert np.array_equal(oracle.knn_brute(g["pts"], g["q"], 7), g["ids"])
    ids, _, _, _ = oracle.kdtree_knn(g["pts"], g["q"], 7)
    assert np.array_equal(ids, g["ids"])


def _np_knn(pts, q, k):
    """Tiny-input brute force: numpy float64 distances + lexsort on (d, id)."""
    dx = q[:, None, 0].astype(np.float64) - pts[None, :, 0].astype(np.float64)
    dy = q[:, None, 1].astype(np.float64) - pts[None, :, 1].astype(np.float64)
    dz = q[:, None, 2].astype(np.float64) - pts[None, :, 2].astype(np.float64)
    d = (dx * dx + dy * dy) + dz * dz
    out = np.full((len(q), k), 0xFFFFFFFF, np.uint32)
    dout = np.full((len(q), k), np.inf)
    for r in range(len(q)):
        o = np.lexsort((np.arange(len(pts)), d[r]))[:k]
        out[r, : len(o)] = o
        dout[r, : len(o)] = d[r, o]
    return out, dout


@pytest.mark.parametrize("n,k", [(1, 1), (3, 5), (50, 5), (200, 16), (700, 64)])
def test_knn_tiny_brute_force(n, k):
    rng = np.random.default_rng(n * 100 + k)
    pts = rng.normal(size=(n, 3)).astype(np.float32)
    pts[n // 2] = pts[0]  # duplicate point
    q = np.concatenate([pts[:5], rng.normal(size=(20, 3)).astype(np.float32)])
    ids, d2 = oracle.knn_brute(pts, q, k, with_d2=True)
    ref, dref = _np_knn(pts, q, k)
    assert np.array_equal(ids, ref)
    assert np.array_equal(d2, dref)


@pytest.mark.parametrize("dist", ["uniform", "clustered", "lattice", "duplicates", "road_scale"])
def test_kdtree_matches_brute(dist):
    rng = np.random.default_rng(7)
    if dist == "uniform":
        pts = rng.random((20000, 3), dtype=np.float32)
    elif dist == "clustered":
        pts = (rng.normal(size=(20000, 3)) * np.where(rng.random((20000, 1)) < 0.5, 0.01, 1.0)).astype(np.float32)
    elif dist == "lattice":
        g = np.arange(20, dtype=np.float32)
        pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).copy()
    elif dist == "duplicates":
        pts = np.repeat(rng.random((500, 3), dtype=np.float32), 20, axis=0)
    else:
        pts = (rng.random((20000, 3)) * [1e5, 1e5, 200]).astype(np.float32)
    q = np.concatenate([pts[rng.integers(0, len(pts), 500)], (pts[rng.integers(0, len(pts), 500)] + 0.5)]).astype(
        np.float32)
    for k in (1, 5, 32):
        a = oracle.knn_brute(pts, q, k)
        b, _, _, _ = oracle.kdtree_knn(pts, q, k)
        assert np.array_equal(a, b), (dist, k)


def test_knn_checker_catches_planted_mistakes():
    """The comparisons above are not vacuous: plausible mistakes in an
    implementation (dropped tie-break, swapped neighbours, a dropped axis)
    produce outputs that differ from the oracle."""
    g = np.load(os.path.join(GOLD, "knn_lattice.npz"))
    good = oracle.knn_brute(g["pts"], g["q"], 7)
    # descending-id tie-break
    d = ((g["q"][:, None, :].astype(np.float64) - g["pts"][None].astype(np.float64)) ** 2).sum(-1)
    bad = np.stack([np.lexsort((-np.arange(len(g["pts"])), row))[:7] for row in d]).astype(np.uint32)
    assert not np.array_equal(bad, good)
    # two neighbours swapped
    sw = good.copy()
    sw[:, [0, 1]] = sw[:, [1, 0]]
    assert not np.array_equal(sw, good)
    # z axis dropped from the distance
    pts, q = g["pts"].copy(), g["q"].copy()
    pts2, q2 = pts.copy(), q.copy()
    pts2[:, 2] = 0
    q2[:, 2] = 0
    assert not np.array_equal(oracle.knn_brute(pts2, q2, 7), good)


# ---------------------------------------------------------------- range query
def test_range_matches_numpy_and_scipy():
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(8)
    pts = rng.random((5000, 3), dtype=np.float32)
    q = rng.random((300, 3), dtype=np.float32)
    r = 0.05
    off, ids = oracle.range_brute(pts, q, r)
    d = ((q[:, None, :].astype(np.float64) - pts[None].astype(np.float64)) ** 2)
    d = (d[..., 0] + d[..., 1]) + d[..., 2]
    r2 = float(np.float64(np.float32(r)) ** 2)
    tree = cKDTree(pts.astype(np.float64))
    for i in range(len(q)):
        got = ids[off[i]: off[i + 1]]
        want = np.nonzero(d[i] <= r2)[0]
        want = want[np.lexsort((want, d[i, want]))]
        assert np.array_equal(got, want)
        sc = set(tree.query_ball_point(q[i].astype(np.float64), float(np.float32(r))))
        near = np.abs(np.sqrt(d[i]) - np.float32(r)) < 1e-9
        assert set(got.tolist()) - set(np.nonzero(near)[0]) == sc - set(np.nonzero(near)[0])


def test_range_boundary_inclusive():
    pts = np.array([[0, 0, 0], [1, 0, 0], [0, 2, 0], [3, 4, 0]], np.float32)
    q = np.zeros((1, 3), np.float32)
    off, ids = oracle.range_brute(pts, q, 5.0)  # |(3,4,0)| == 5 exactly: included
    assert ids.tolist() == [0, 1, 2, 3]
    off, ids = oracle.range_brute(pts, q, 2.0)
    assert ids.tolist() == [0, 1, 2]


# ---------------------------------------------------------------- Barnes-Hut reference
def test_bh_direct_two_body_closed_form():
    pm = np.array([[0, 0, 0, 2.0], [3, 4, 0, 5.0]], np.float32)
    acc = oracle.bh_direct(pm, G=1.5, eps=0.5)
    r2 = 25.0 + 0.25
    want0 = 1.5 * 5.0 * np.array([3, 4, 0]) / r2 ** 1.5
    want1 = -1.5 * 2.0 * np.array([3, 4, 0]) / r2 ** 1.5
    assert np.allclose(acc[0], want0, rtol=1e-14)
    assert np.allclose(acc[1], want1, rtol=1e-14)


def test_bh_direct_matches_numpy_and_conserves_momentum():
    rng = np.random.default_rng(9)
    pm = np.concatenate([rng.normal(size=(400, 3)), rng.uniform(0.5, 1.5, (400, 1))], 1).astype(np.float32)
    acc = oracle.bh_direct(pm, G=1.0, eps=1e-3)
    x = pm[:, :3].astype(np.float64)
    m = pm[:, 3].astype(np.float64)
    diff = x[None, :, :] - x[:, None, :]
    r2 = (diff ** 2).sum(-1) + 1e-6
    want = (m[None, :, None] * diff / r2[..., None] ** 1.5).sum(1)
    assert np.allclose(acc, want, rtol=1e-10, atol=1e-12)
    assert np.abs((m[:, None] * acc).sum(0)).max() < 1e-9 * np.abs(m[:, None] * acc).sum()
    sample = np.array([5, 17, 399], np.int64)
    assert np.array_equal(oracle.bh_direct(pm, sample, G=1.0, eps=1e-3), acc[sample])


def test_bh_direct_reference_law_properties():
    """The reference shader's law, G m_j d / (|d|^2 sqrt(|d|^2 + eps^2)) over
    pairs with |d|^2 > 1e-10: equal to the Plummer law far away, exactly twice
    it at |d| = eps, pairs inside the cut-off contribute nothing (coincident
    bodies included), and forces are antisymmetric (momentum is conserved)."""
    eps = 0.5
    far = np.array([[0, 0, 0, 2.0], [600, 800, 0, 5.0]], np.float32)
    a_ref = oracle.bh_direct(far, G=1.5, eps=eps, law="reference")
    a_pl = oracle.bh_direct(far, G=1.5, eps=eps)
    assert np.allclose(a_ref, a_pl, rtol=1e-6) and not np.array_equal(a_ref, a_pl)
    assert np.allclose(a_ref[0], 1.5 * 5.0 * np.array([0.6, 0.8, 0]) / 1000.0 ** 2, rtol=1e-6)
    at_eps = np.array([[0, 0, 0, 1.0], [0.5, 0, 0, 3.0]], np.float32)
    ratio = oracle.bh_direct(at_eps, eps=eps, law="reference")[:, 0] / oracle.bh_direct(at_eps, eps=eps)[:, 0]
    assert np.allclose(ratio, 2.0, rtol=1e-12)
    # cut-off: d2 = 0 and d2 = 1e-12 skipped, d2 = 4e-10 kept
    pm = np.array([[1, 1, 1, 1.0], [1, 1, 1, 2.0], [1 + 1e-6, 1, 1, 4.0]], np.float32)
    d = np.float64(np.float32(1 + 1e-6)) - 1.0
    acc = oracle.bh_direct(pm, G=1.0, eps=1e-3, law="reference")
    assert np.all(np.isfinite(acc)) and np.all(acc[:, 1:] == 0)
    assert d * d < 1e-10 and np.all(acc == 0)
    pm[2, 0] = np.float32(1 + 2e-5)
    d = np.float64(pm[2, 0]) - 1.0
    acc = oracle.bh_direct(pm, G=1.0, eps=1e-3, law="reference")
    assert d * d > 1e-10 and np.isclose(acc[0, 0], 4.0 * d / (d * d * np.sqrt(d * d + 1e-6)), rtol=1e-9)
    rng = np.random.default_rng(10)
    pm = np.concatenate([rng.normal(size=(300, 3)), rng.uniform(0.5, 1.5, (300, 1))], 1).astype(np.float32)
    acc = oracle.bh_direct(pm, G=1.0, eps=1e-3, law="reference")
    m = pm[:, 3].astype(np.float64)
    assert np.abs((m[:, None] * acc).sum(0)).max() < 1e-9 * np.abs(m[:, None] * acc).sum()
    with pytest.raises(ValueError):
        oracle.bh_direct(pm, law="newton")


def test_d2_contract_is_binary64_without_fma():
    q = np.array([0.1, 0.2, 0.3], np.float32)
    p = np.array([0.7, -0.4, 1e-3], np.float32)
    dx, dy, dz = (np.float64(q[i]) - np.float64(p[i]) for i in range(3))
    assert oracle.lib().oracle_d2_64(q, p) == (dx * dx + dy * dy) + dz * dz
    assert math.isfinite(oracle.lib().oracle_d2_64(q, p))
