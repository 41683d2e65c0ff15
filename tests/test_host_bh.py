"""The Barnes-Hut walk (bh_walk.hpp) and the mass refit built for the host
(libgprt_host.so), against the FP64 direct sum of the oracle, on the CPU.

Checks, none fitted to observed errors (u = 2^-24):
- the FP32 sums lie within ROUND_U * u * S of the same accepted children and
  bodies summed in FP64, S being the sum of the terms' magnitudes.  First-order
  rounding count: the offset (u), the fma chain of |d|^2 + eps^2 (5 u with the
  squared offsets), sqrt and the division (2 u), the cube and the products with
  G m (5 u), inherited by the weight (about 15.5 u from r^2) and the product with
  the offset (2 u), about 21 u per term; the Neumaier sum adds 2 u (plus
  O(n u^2)), the final s + c one more; the FP32 eps^2 against the FP64 one 1.5 u;
- Plummer law: the FP64 sum lies within the derived monopole bound
  (host_walk.cpp: Taylor remainder of the softened kernel about each accepted
  child's FP32 centre of mass, with the centre and mass rounding terms) of the
  direct sum, plus the FP64 rounding of two sums in different orders;
- theta = 0 accepts no child: the bound is zero and the sums agree to rounding.
A planted mistake in the direct sum (G off by 25%) must break the bound."""
import numpy as np
import pytest

import oracle
from gprt_hip import datasets as D
from gprt_hip import hostwalk as H

U = 2.0 ** -24
ROUND_U = 32.0
F64_SLACK = 1e-10  # two FP64 sums of up to 1e6 terms in different orders, relative to S
EPS = float(np.float32(1e-3))  # the kernel's eps, so the direct sum uses the same value


def bodies(kind, n, rng):
    """Seeded body sets: Plummer sphere, uniform cube, tight clusters, exact
    duplicates, a plane; masses uniform in [0.5, 1.5] / n."""
    if kind == "plummer":
        return D.plummer(n, seed=int(rng.integers(1 << 30)))
    if kind == "uniform":
        xyz = rng.uniform(-100, 100, (n, 3))
    elif kind == "clusters":
        c = rng.normal(size=(5, 3)) * 10
        xyz = c[rng.integers(0, 5, n)] + rng.normal(size=(n, 3)) * 0.01
    elif kind == "duplicates":
        base = rng.normal(size=(max(n // 4, 1), 3))
        xyz = base[rng.integers(0, len(base), n)]
    else:  # plane
        xyz = np.c_[rng.uniform(-1, 1, (n, 2)), np.zeros(n)]
    m = rng.uniform(0.5, 1.5, (n, 1)) / n
    return np.ascontiguousarray(np.c_[xyz, m], np.float32)


def rounding_violations(acc, h):
    """Bodies whose FP32 result (acc, float) is further than ROUND_U u S from
    the FP64 sum of the same terms."""
    err = np.abs(np.asarray(acc, np.float64) - h["acc64"]).max(axis=1)
    return np.nonzero(err > ROUND_U * U * h["absum"])[0]


def bound_violations(acc, h, direct):
    """Bodies whose result is further from the direct sum than the derived
    monopole bound (which carries |G|) plus the rounding allowances."""
    err = np.abs(np.asarray(acc, np.float64) - direct).max(axis=1)
    lim = h["bound"] + (ROUND_U * U + F64_SLACK) * h["absum"]
    return np.nonzero(~(err <= lim))[0]


CASES = [("plummer", 2000, 8), ("uniform", 3000, 1), ("clusters", 3000, 4), ("duplicates", 1000, 8),
         ("plane", 2000, 32), ("plummer", 20000, 8)]


@pytest.mark.parametrize("kind,n,leaf", CASES)
@pytest.mark.parametrize("theta", [0.0, 0.3, 0.5, 0.7])
def test_host_bh_within_derived_bound(kind, n, leaf, theta):
    rng = np.random.default_rng(n + leaf)
    pm = bodies(kind, n, rng)
    sample = np.unique(np.r_[np.arange(min(n, 8)), rng.integers(0, n, 400)])
    h = H.barnes_hut(pm, theta, eps=EPS, leaf_size=leaf, sample=sample)
    direct = oracle.bh_direct(pm, sample, G=1.0, eps=EPS)
    assert len(rounding_violations(h["acc32"], h)) == 0
    if theta == 0.0:
        assert h["stats"][2] == 0 and np.all(h["bound"] == 0.0)
    bad = bound_violations(h["acc64"], h, direct)
    assert len(bad) == 0, f"{kind} n={n} theta={theta}: {len(bad)} bodies outside the bound"
    bad32 = bound_violations(h["acc32"], h, direct)
    assert len(bad32) == 0


def test_host_bh_reference_law_rounding():
    """The reference force law has no derived bound here; its FP32 sums stay
    within the rounding allowance of the FP64 ones, and at theta = 0 equal the
    direct sum (pairs with |d|^2 <= 1e-10 skipped, duplicates included)."""
    rng = np.random.default_rng(7)
    for kind in ("plummer", "duplicates"):
        pm = bodies(kind, 2000, rng)
        for theta in (0.0, 0.5):
            h = H.barnes_hut(pm, theta, eps=EPS, law="reference")
            assert len(rounding_violations(h["acc32"], h)) == 0
            assert np.all(np.isnan(h["bound"]))
        direct = oracle.bh_direct(pm, None, G=1.0, eps=EPS, law="reference")
        h0 = H.barnes_hut(pm, 0.0, eps=EPS, law="reference")
        err = np.abs(h0["acc64"] - direct).max(axis=1)
        assert np.all(err <= F64_SLACK * h0["absum"])


def test_planted_mistake_breaks_the_bound():
    """The bound is not vacuous: a direct sum computed with G = 1.25 leaves
    most bodies outside it at theta = 0.3."""
    rng = np.random.default_rng(11)
    pm = bodies("plummer", 5000, rng)
    sample = np.arange(0, 5000, 10)
    h = H.barnes_hut(pm, 0.3, eps=EPS, sample=sample)
    wrong = oracle.bh_direct(pm, sample, G=1.25, eps=EPS)
    assert len(bound_violations(h["acc64"], h, wrong)) > len(sample) // 2


def test_host_bh_counts_and_limits():
    """One body; two bodies under the unsoftened reference law (exact); bad
    arguments; counts (theta = 0 sums every pair body by body)."""
    pm = np.array([[0, 0, 0, 1]], np.float32)
    h = H.barnes_hut(pm, 0.5)
    assert np.all(h["acc32"] == 0) and h["stats"][1] == 1
    pm = np.array([[0, 0, 0, 1], [1, 0, 0, 2]], np.float32)
    h = H.barnes_hut(pm, 0.5, eps=0.0, leaf_size=1, law="reference")
    np.testing.assert_allclose(h["acc64"], [[2, 0, 0], [-1, 0, 0]], rtol=1e-15)
    assert np.array_equal(h["acc32"], np.float32([[2, 0, 0], [-1, 0, 0]]))
    with pytest.raises(RuntimeError):
        H.barnes_hut(pm, 0.5, sample=[2])
    with pytest.raises(RuntimeError, match="eps must be > 0"):
        H.barnes_hut(pm, 0.5, eps=0.0)
    rng = np.random.default_rng(2)
    pm = bodies("uniform", 3000, rng)
    for theta in (0.0, 0.5, 1.0):
        h = H.barnes_hut(pm, theta, leaf_size=4)
        v, b, a = h["stats"]
        if theta == 0.0:
            assert b == 3000 * 3000 and a == 0
        else:
            assert b < 3000 * 3000 and 0 < a <= 2 * v
