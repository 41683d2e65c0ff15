"""A seeded matrix of k-NN cases on the GPU against the FP64 brute-force oracle:
point distributions (uniform, Gaussian, clustered, a plane, an integer lattice
with exact ties, a thin line), cloud sizes from 1 to 60,000, k across every
candidate class, and the schedule knobs (traversal, LDS ring, compacted rounds,
radius heuristic and multiplier, graph replay).  Every case must return the
oracle's ids bit for bit, in both forms."""
import numpy as np
import pytest

import gprt_hip
import oracle

pytestmark = pytest.mark.gpu

KS = [1, 3, 5, 8, 12, 16, 21, 24, 29, 37, 44, 55, 61, 64]


def _cloud(kind, n, rng):
    if kind == "uniform":
        return rng.random((n, 3), dtype=np.float32)
    if kind == "gauss":
        return (rng.normal(size=(n, 3)) * 3.0).astype(np.float32)
    if kind == "clusters":
        c = rng.random((8, 3)) * 100.0
        return (c[rng.integers(0, 8, n)] + rng.normal(size=(n, 3)) * 0.05).astype(np.float32)
    if kind == "plane":
        p = rng.random((n, 3), dtype=np.float32) * 50.0
        p[:, 2] = 1.25
        return p
    if kind == "lattice":
        g = int(np.ceil(n ** (1 / 3)))
        a = np.arange(g, dtype=np.float32)
        p = np.stack(np.meshgrid(a, a, a, indexing="ij"), -1).reshape(-1, 3)
        return np.ascontiguousarray(p[:n])
    # "line": points along x with tiny jitter in y, z
    p = np.zeros((n, 3), np.float32)
    p[:, 0] = rng.random(n, dtype=np.float32) * 1e3
    p[:, 1:] = rng.normal(size=(n, 2)).astype(np.float32) * 1e-3
    return p


def _queries(pts, nq, rng):
    if len(pts) == 0:
        return rng.random((nq, 3), dtype=np.float32)
    base = pts[rng.integers(0, len(pts), nq)]
    span = float(np.ptp(pts, axis=0).max()) if len(pts) > 1 else 1.0
    jitter = rng.normal(size=(nq, 3)).astype(np.float32) * np.float32(0.01 * max(span, 1e-3))
    q = (base + jitter).astype(np.float32)
    q[: nq // 10] = base[: nq // 10]  # exact self-matches
    return q


CONFIGS = [
    dict(),
    dict(traversal=gprt_hip.TRAV_WIDE),
    dict(lds_stack=16, stack_depth=20),
    dict(compact_rounds=1, radius_heuristic=gprt_hip.RADIUS_FIXED, radius_param=1e-3, max_rounds=6),
    dict(radius_heuristic=gprt_hip.RADIUS_INFINITE, use_graph=0),
    dict(radius_param=0.1, radius_mult=1.5, sort_queries=0),
    dict(traversal=gprt_hip.TRAV_WAVE),
    dict(sampler=gprt_hip.SAMPLER_STRIDED, radius_param=0.5),
]

CASES = []
_kinds = ["uniform", "gauss", "clusters", "plane", "lattice", "line"]
_sizes = [1, 7, 300, 5_000, 60_000]
for _i in range(96):
    CASES.append((_i, _kinds[_i % len(_kinds)], _sizes[(_i * 7) % len(_sizes)], KS[(_i * 5) % len(KS)],
                  _i % len(CONFIGS)))
# k beyond 64 (one query per wavefront): every distribution, sizes around k
for _i in range(24):
    CASES.append((200 + _i, _kinds[_i % len(_kinds)], [1, 90, 300, 5_000][_i % 4], [65, 100, 200, 777][(_i * 3) % 4],
                  _i % len(CONFIGS)))


@pytest.mark.parametrize("seed,kind,n,k,ci", CASES)
def test_parity_matrix(ctx, seed, kind, n, k, ci):
    rng = np.random.default_rng(1000 + seed)
    pts = _cloud(kind, n, rng)
    q = _queries(pts, 1_500, rng)
    ref = oracle.knn_brute(pts, q, k)
    leaf = (4, 8, 16)[seed % 3]
    t = ctx.build(pts, leaf_size=leaf)
    cfg = gprt_hip.default_config(**CONFIGS[ci])
    got = t.k_nn(q, k, cfg)
    bad = np.nonzero((got != ref).any(axis=1))[0]
    assert len(bad) == 0, f"{kind} n={n} k={k} cfg={CONFIGS[ci]}: {len(bad)} queries differ, first {bad[:3]}"
    assert np.array_equal(t.k_nn_plain(q, k), ref)
    t.close()


RANGE_CASES = []
for _i in range(30):
    RANGE_CASES.append((_i, _kinds[_i % len(_kinds)], [1, 300, 5_000, 40_000][(_i * 3) % 4],
                        [0.0, 1e-3, 0.01, 0.05, 0.2][(_i * 7) % 5]))


@pytest.mark.parametrize("seed,kind,n,rel", RANGE_CASES)
def test_range_parity_matrix(ctx, seed, kind, n, rel):
    """Range queries over the same distributions: every point with
    d64 <= r*r, in (d64, id) order, CSR offsets equal to the oracle's."""
    rng = np.random.default_rng(2000 + seed)
    pts = _cloud(kind, n, rng)
    q = _queries(pts, 800, rng)
    span = float(np.ptp(pts, axis=0).max()) if len(pts) > 1 else 1.0
    r = float(np.float32(rel * max(span, 1e-3)))
    roff, rids = oracle.range_brute(pts, q, r)
    t = ctx.build(pts, leaf_size=(4, 8, 16)[seed % 3])
    off, ids = t.range_query(q, r)
    assert np.array_equal(off, roff), f"{kind} n={n} r={r}: counts differ"
    assert np.array_equal(ids, rids), f"{kind} n={n} r={r}: ids or order differ"
    t.close()
