"""Writes the committed golden fixtures used to pin the oracle.

    python tests/golden/make_golden.py

The reference ships no golden vectors for this path (its Rust/OptiX code has
no tests and cannot run here), so the fixtures come from independent sources:

  knn_uniform.npz   -- 2,000 uniform points, 300 jittered queries (seed 11);
                       exact 5-NN ids and distances from scipy.spatial.cKDTree
                       (a library independent of the oracle).
  knn_lattice.npz   -- 6^3 integer lattice, queries on lattice points / cell
                       centres (exact ties); 7-NN ids by a numpy lexsort over
                       (d64, id) computed here with float64 arrays.
  lcg_draws.json    -- the first 32 draws (seed >> 33) of the LCG of
                       gprt-optix/src/dispatch.rs:65-69 from seed
                       0x123456789ABCDEF, in Python integer arithmetic.
  morton_known.json -- hand-derivable Morton codes (bit interleave by hand).
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def knn_uniform():
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(11)
    pts = rng.random((2000, 3), dtype=np.float32)
    q = (pts[rng.integers(0, 2000, 300)] + rng.normal(0, 1e-2, (300, 3))).astype(np.float32)
    d, i = cKDTree(pts.astype(np.float64)).query(q.astype(np.float64), 5)
    np.savez_compressed(os.path.join(HERE, "knn_uniform.npz"), pts=pts, q=q, ids=i.astype(np.uint32),
                        dist=d.astype(np.float64))


def knn_lattice():
    g = np.arange(6, dtype=np.float32)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).copy()
    rng = np.random.default_rng(12)
    q = np.concatenate([pts[rng.integers(0, len(pts), 40)], pts[rng.integers(0, len(pts), 40)] + 0.5]).astype(
        np.float32)
    d = ((q[:, None, :].astype(np.float64) - pts[None, :, :].astype(np.float64)) ** 2).sum(-1)
    ids = np.empty((len(q), 7), np.uint32)
    for r in range(len(q)):
        order = np.lexsort((np.arange(len(pts)), d[r]))
        ids[r] = order[:7]
    np.savez_compressed(os.path.join(HERE, "knn_lattice.npz"), pts=pts, q=q, ids=ids)


def lcg_draws():
    s = 0x123456789ABCDEF
    out = []
    for _ in range(32):
        s = (s * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        out.append(s >> 33)
    with open(os.path.join(HERE, "lcg_draws.json"), "w") as f:
        json.dump({"seed": "0x123456789ABCDEF", "draws": out}, f, indent=1)


def morton_known():
    # x bits go to positions 3i+2, y to 3i+1, z to 3i
    known = [
        [[1, 0, 0], 4], [[0, 1, 0], 2], [[0, 0, 1], 1], [[1, 1, 1], 7],
        [[2, 0, 0], 32], [[0, 2, 0], 16], [[0, 0, 2], 8],
        [[1023, 1023, 1023], 0x3FFFFFFF], [[1023, 0, 0], 0x24924924], [[0, 1023, 0], 0x12492492],
        [[0, 0, 1023], 0x09249249], [[512, 0, 0], 1 << 29], [[3, 5, 6], 0b011_101_110 ^ 0],
    ]
    with open(os.path.join(HERE, "morton_known.json"), "w") as f:
        json.dump(known, f, indent=1)


if __name__ == "__main__":
    knn_uniform()
    knn_lattice()
    lcg_draws()
    morton_known()
    print("fixtures written to", HERE)
