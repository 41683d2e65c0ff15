"""Parity at the benchmark shapes themselves, one GPU:

  road network   434,000 points x all 400,000 queries, k = 5, both forms
  LiDAR range    r = 0.5 m over the full 1,310,720-point scan, all 400,000
                 queries (557.6 M hits): 2,000 sampled queries exactly against
                 FP64 brute force, every segment's structure and a 20,000-query
                 sample's (d64, id) order checked over the whole output
  256 clouds     every cloud of the uneven batch (1 K .. 200 K points, k in
                 {1, 5, 16, 64}) against its own KD-tree, one call
  Barnes-Hut     1,000,000-body Plummer sphere at theta 0.3 / 0.5 / 0.7 on
                 10,000 sampled bodies: the host build bit for bit, and the
                 derived bound against the FP64 direct sum

(The LiDAR k = 5 and k = 32 k-NN at full size is test_gpu_knn.py::test_knn_lidar_full.)"""
import numpy as np
import pytest

import gprt_hip
import oracle
from gprt_hip import datasets as D
from gprt_hip import hostwalk as H
from test_host_bh import EPS, bound_violations

pytestmark = pytest.mark.gpu


def test_road_network_full(ctx):
    pts, q = D.road_network()
    assert (len(pts), len(q)) == (434_000, 400_000)
    ref, _, _, _ = oracle.kdtree_knn(pts, q, 5)
    t = ctx.build(pts)
    try:
        for label, got in (("batched", t.k_nn(q, 5)), ("plain", t.k_nn_plain(q, 5))):
            bad = np.nonzero((got != ref).any(axis=1))[0]
            assert len(bad) == 0, f"road {label}: {len(bad)} of 400,000 queries differ, first {bad[:5]}"
    finally:
        t.close()


def test_lidar_range_full(ctx):
    pts, q = D.lidar_scan()
    r = 0.5
    t = ctx.build(pts)
    try:
        off, ids = t.range_query(q, r)
    finally:
        t.close()
    n = len(pts)
    assert off[0] == 0 and np.all(np.diff(off) >= 0) and off[-1] == len(ids)
    print(f"lidar range r=0.5: {off[-1]} hits, mean {off[-1] / len(q):.1f}, max {np.diff(off).max()}")
    assert ids.max() < n
    rng = np.random.default_rng(21)
    sel = np.sort(rng.choice(len(q), 2_000, replace=False))
    roff, rids = oracle.range_brute(pts, q[sel], r)
    assert np.array_equal(np.diff(off)[sel], np.diff(roff)), "hit counts differ from FP64 brute force"
    got = np.concatenate([ids[off[i]:off[i + 1]] for i in sel])
    assert np.array_equal(got, rids), "ids or their (d64, id) order differ from FP64 brute force"
    # the whole output's structure on a larger sample: every hit within r in
    # binary64, no id twice in a segment, segments ascending by (d64, id)
    r2 = np.float64(np.float32(r)) ** 2
    for i in np.sort(rng.choice(len(q), 20_000, replace=False)):
        seg = ids[off[i]:off[i + 1]].astype(np.int64)
        if len(seg) == 0:
            continue
        d = pts[seg].astype(np.float64) - q[i].astype(np.float64)
        d64 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        assert np.all(d64 <= r2)
        assert len(np.unique(seg)) == len(seg)
        ok = (d64[1:] > d64[:-1]) | ((d64[1:] == d64[:-1]) & (seg[1:] > seg[:-1]))
        assert ok.all(), f"query {i}: segment not in (d64, id) order"


def test_batched_clouds_all(ctx):
    pts, poff, q, qoff, kc = D.clustered_batch()
    assert len(kc) == 256
    t = ctx.build_batched(pts, poff, leaf_size=8)
    try:
        ids, out_off = t.k_nn_batched(q, qoff, kc, gprt_hip.default_config())
    finally:
        t.close()
    bad = []
    for c in range(len(kc)):
        k = int(kc[c])
        ref, _, _, _ = oracle.kdtree_knn(pts[poff[c]:poff[c + 1]], q[qoff[c]:qoff[c + 1]], k)
        got = ids[out_off[c]:out_off[c + 1]].reshape(-1, k)
        if not np.array_equal(got, ref):
            bad.append((c, k, int(poff[c + 1] - poff[c])))
    print(f"256 clouds: {len(pts)} points, {len(q)} queries")
    assert not bad, f"clouds differing from their KD-tree (cloud, k, n): {bad[:10]}"


def test_plummer_1m_errors(ctx):
    """1,000,000-body Plummer sphere at theta 0.3 / 0.5 / 0.7 on 10,000 sampled
    bodies: the kernel equals the host build of the same walk bit for bit, lies
    within the derived monopole bound of the FP64 direct sum (test_host_bh.py),
    and passes the reference's own physics check, total |F| over the sample
    within 2% of the direct sum (gprt-sandbox/verify_physics.py:5-48).  The
    relative errors are printed per theta."""
    pm = D.plummer()
    sample = np.sort(np.random.default_rng(5).choice(len(pm), 10_000, replace=False))
    ref = oracle.bh_direct(pm, sample, G=1.0, eps=EPS)
    dpm = gprt_hip.DeviceArray.from_host(ctx, pm)
    try:
        for theta in (0.3, 0.5, 0.7):
            acc = ctx.barnes_hut(dpm, theta=theta, eps=EPS, G=1.0)
            a = acc.to_host()[sample]
            acc.free()
            assert np.all(np.isfinite(a))
            h = H.barnes_hut(pm, theta, eps=EPS, G=1.0, sample=sample)
            diff = np.nonzero((a.view(np.uint32) != h["acc32"].view(np.uint32)).any(axis=1))[0]
            assert len(diff) == 0, f"theta={theta}: {len(diff)} bodies differ from the host build"
            bad = bound_violations(a, h, ref)
            assert len(bad) == 0, f"theta={theta}: {len(bad)} bodies outside the derived bound"
            a = a.astype(np.float64)
            err = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
            tot = abs(np.linalg.norm(a, axis=1).sum() / np.linalg.norm(ref, axis=1).sum() - 1.0)
            print(f"plummer 1M theta={theta}: median {np.median(err):.2e} p99 {np.quantile(err, 0.99):.2e} "
                  f"max {err.max():.2e} total |F| {tot:.2e}; bound/|a| median "
                  f"{np.median(h['bound'] / np.linalg.norm(ref, axis=1)):.2e}")
            assert tot < 0.02
    finally:
        dpm.free()
