"""A seeded matrix of Barnes-Hut cases on the GPU: body distributions
(Plummer, uniform, tight clusters, exact duplicates, a plane), sizes from 1 to
20,000, leaf sizes 1 to 32, both force laws, theta 0, 0.5 and 1.

Per case, none of the checks fitted to observed errors:
- the kernel's accelerations equal, bit for bit, those of the host build of
  the same walk and mass refit (gprt_hip.hostwalk.barnes_hut), and the
  wavefront walk (the default) equals the per-lane walk for every body;
- they lie within the rounding allowance of the same terms summed in FP64;
- Plummer law: they lie within the derived monopole bound of the FP64 direct
  sum plus that allowance (test_host_bh.py states both); at theta = 0 the bound
  is zero;
- reference law at theta = 0: the FP64 direct sum to the rounding allowance."""
import numpy as np
import pytest

import gprt_hip
import oracle
from gprt_hip import hostwalk as H
from test_host_bh import EPS, F64_SLACK, ROUND_U, U, bodies, bound_violations, rounding_violations

pytestmark = pytest.mark.gpu

KINDS = ["plummer", "uniform", "clusters", "duplicates", "plane"]
SIZES = [1, 2, 33, 1_000, 20_000]
CASES = [(i, KINDS[i % 5], SIZES[(i * 3) % 5], (1, 4, 8, 32)[i % 4], ("plummer", "reference")[(i // 2) % 2])
         for i in range(20)]


@pytest.mark.parametrize("seed,kind,n,leaf,law", CASES)
def test_bh_matrix(ctx, seed, kind, n, leaf, law):
    rng = np.random.default_rng(3000 + seed)
    pm = bodies(kind, n, rng)
    sample = np.unique(np.r_[np.arange(min(n, 8)), rng.integers(0, n, 200)])
    direct = oracle.bh_direct(pm, sample, G=1.0, eps=EPS, law=law)
    for theta in (0.0, 0.5, 1.0):
        acc = ctx.barnes_hut(pm, theta=theta, eps=EPS, G=1.0, leaf_size=leaf, law=law)
        assert np.all(np.isfinite(acc))
        # the per-lane walk gives the wavefront walk's bits for every body
        lane = ctx.barnes_hut(pm, theta=theta, eps=EPS, G=1.0, leaf_size=leaf, law=law,
                              flags=gprt_hip.BH_LANE_WALK)
        assert np.array_equal(acc.view(np.uint32), lane.view(np.uint32)), (kind, n, leaf, law, theta)
        h = H.barnes_hut(pm, theta, eps=EPS, G=1.0, leaf_size=leaf, law=law, sample=sample)
        a = acc[sample]
        diff = np.nonzero((a.view(np.uint32) != h["acc32"].view(np.uint32)).any(axis=1))[0]
        assert len(diff) == 0, (f"{kind} n={n} leaf={leaf} {law} theta={theta}: {len(diff)} bodies differ from "
                                f"the host build, first {sample[diff[:5]]}")
        assert len(rounding_violations(a, h)) == 0
        if law == "plummer":
            bad = bound_violations(a, h, direct)
            assert len(bad) == 0, f"{kind} n={n} leaf={leaf} theta={theta}: bodies {sample[bad[:5]]} outside the bound"
        elif theta == 0.0:
            err = np.abs(a.astype(np.float64) - direct).max(axis=1)
            assert np.all(err <= (ROUND_U * U + F64_SLACK) * h["absum"])


def test_bh_counters_match_host(ctx):
    """Node visits, bodies summed and children accepted, counted on the GPU
    over every body, equal the host build's counts."""
    rng = np.random.default_rng(77)
    pm = bodies("plummer", 20_000, rng)
    for theta in (0.3, 0.7):
        h = H.barnes_hut(pm, theta, eps=EPS)
        for walk in (0, gprt_hip.BH_LANE_WALK):
            ctx.barnes_hut(pm, theta=theta, eps=EPS, flags=gprt_hip.COLLECT_COUNTERS | walk)
            s = ctx.stats()
            assert (s["node_visits"], s["point_tests"], s["rerank_reads"]) == tuple(int(x) for x in h["stats"])
            if walk == 0:
                # a wavefront loads a node (a bucket) once for all of its lanes that visit it
                assert 0 < s["wave_steps"] <= s["node_visits"]
                assert 0 < s["lane_steps"] <= s["point_tests"]


def test_bh_plummer_eps_zero_rejected(ctx):
    pm = np.array([[0, 0, 0, 1], [1, 0, 0, 2]], np.float32)
    with pytest.raises(gprt_hip.GprtError, match="eps must be > 0"):
        ctx.barnes_hut(pm, theta=0.5, eps=0.0)
    acc = ctx.barnes_hut(pm, theta=0.5, eps=0.0, leaf_size=1, law="reference")
    assert np.array_equal(acc, np.float32([[2, 0, 0], [-1, 0, 0]]))
