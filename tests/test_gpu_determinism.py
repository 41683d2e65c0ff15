"""Determinism: the same inputs give the same bytes, call after call and
across schedules that must not change results (captured graph or not, compact
rounds or inline ones), for every product entry point: k-NN ids and FP64
distances (batched, plain and multi-cloud), range CSR, Barnes-Hut
accelerations (raw float bits), and the tree itself (exported keys,
permutation and nodes of two builds of the same points)."""
import numpy as np
import pytest

import gprt_hip
from gprt_hip import datasets as D

pytestmark = pytest.mark.gpu


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint8).tobytes()


def test_knn_and_range_bytes_repeat(ctx):
    pts, q = D.lidar_scan(frames=3, nq=20_000)
    t = ctx.build(pts)
    runs = []
    for cfg in (gprt_hip.default_config(), gprt_hip.default_config(use_graph=0),
                gprt_hip.default_config(compact_rounds=1, use_graph=0), gprt_hip.default_config()):
        ids, d2 = t.k_nn(q, 7, cfg, with_d2=True)
        runs.append((_bits(ids), _bits(d2)))
    assert all(r == runs[0] for r in runs[1:])
    pids, pd2 = t.k_nn_plain(q, 7, with_d2=True)
    assert (_bits(pids), _bits(pd2)) == runs[0]
    r1 = t.range_query(q[:3_000], 0.3)
    r2 = t.range_query(q[:3_000], 0.3)
    assert _bits(r1[0]) == _bits(r2[0]) and _bits(r1[1]) == _bits(r2[1])
    t.close()


def test_tree_build_bytes_repeat(ctx):
    pts, _ = D.lidar_scan(frames=2, nq=10)
    exports = []
    for _ in range(2):
        t = ctx.build(pts)
        exports.append(tuple(_bits(a) for a in t.export()))
        t.close()
    assert exports[0] == exports[1]


def test_barnes_hut_and_batched_bytes_repeat(ctx):
    pm = D.plummer(50_000, seed=11)
    a = [_bits(ctx.barnes_hut(pm, theta=0.5, eps=1e-3, G=1.0, law=law)) for law in ("plummer", "plummer",
                                                                                  "reference", "reference")]
    assert a[0] == a[1] and a[2] == a[3] and a[0] != a[2]
    pts, poff, q, qoff, kc = D.clustered_batch(n_clouds=24, lo=1e3, hi=2e4, seed=12)
    t = ctx.build_batched(pts, poff)
    out = [t.k_nn_batched(q, qoff, kc, gprt_hip.default_config(use_graph=g)) for g in (1, 0, 1)]
    assert all(_bits(o[0]) == _bits(out[0][0]) and _bits(o[1]) == _bits(out[0][1]) for o in out[1:])
    t.close()
