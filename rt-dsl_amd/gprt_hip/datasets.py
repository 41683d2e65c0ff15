"""Deterministic synthetic workloads (numpy PCG64 streams, fixed seeds).

The reference's datasets (3DRoad, KITTI, 3DIono, Porto; README.md:67-70) are
not shipped and need a network, so each workload below is generated with the
shape of the corresponding benchmark:

  uniform_cloud   -- 100,000 points uniform in [0,1]^3 + 10,000 jittered queries
  road_network    -- 434,000 points along 5,000 random 2D polylines in a 100 km
                     square, smooth elevation + N(0, 0.5 m); 400,000 queries with
                     N(0, 1 m) jitter                         (3DRoad shape)
  lidar_scan      -- 20 frames x 64 rings x 1,024 azimuths = 1,310,720 points of a
                     ground-plane + boxes + 120 m cylinder scene, N(0, 2 cm) range
                     noise; 400,000 queries with N(0, 2 cm) jitter (KITTI shape)
  clustered_batch -- 256 clouds of log-uniform size in [1e3, 2e5], Gaussian
                     mixtures + 5% background, 10% queries, k in {1, 5, 16, 64}
  plummer         -- Plummer sphere (a = 1, equal masses summing to 1)
"""
import numpy as np


def uniform_cloud(n=100_000, nq=10_000, seed=0, jitter=1e-3):
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 3), dtype=np.float32)
    idx = rng.integers(0, n, nq)
    q = (pts[idx].astype(np.float64) + rng.normal(0.0, jitter, (nq, 3))).astype(np.float32)
    return pts, q


def road_network(n=434_000, nq=400_000, n_lines=5_000, side=100_000.0, seed=2):
    rng = np.random.default_rng(seed)
    nseg = rng.integers(5, 31, n_lines)
    starts = rng.random((n_lines, 2)) * side
    total_seg = int(nseg.sum())
    line_of_seg = np.repeat(np.arange(n_lines), nseg)
    seg_len = rng.uniform(10.0, 500.0, total_seg)
    turn = rng.normal(0.0, 0.3, total_seg)
    head0 = rng.uniform(0, 2 * np.pi, n_lines)
    # heading random walk per polyline
    first = np.concatenate([[0], np.cumsum(nseg)[:-1]])
    turn[first] = head0
    heading = np.cumsum(turn)
    heading -= np.repeat(np.concatenate([[0], np.cumsum(turn)[first[1:] - 1]]), nseg)
    dx = np.cos(heading) * seg_len
    dy = np.sin(heading) * seg_len
    cx = np.cumsum(dx)
    cy = np.cumsum(dy)
    offx = np.repeat(np.concatenate([[0], cx[first[1:] - 1]]), nseg)
    offy = np.repeat(np.concatenate([[0], cy[first[1:] - 1]]), nseg)
    endx = cx - offx + starts[line_of_seg, 0]
    endy = cy - offy + starts[line_of_seg, 1]
    begx = endx - dx
    begy = endy - dy
    # points proportional to segment length
    w = seg_len / seg_len.sum()
    seg_pick = rng.choice(total_seg, size=n, p=w)
    t = rng.random(n)
    x = begx[seg_pick] + t * (endx[seg_pick] - begx[seg_pick])
    y = begy[seg_pick] + t * (endy[seg_pick] - begy[seg_pick])
    x = np.mod(x, side)
    y = np.mod(y, side)

    def elev(x, y):
        return (60.0 * np.sin(2 * np.pi * x / 37_000.0) * np.cos(2 * np.pi * y / 52_000.0)
                + 30.0 * np.sin(2 * np.pi * (x + y) / 13_000.0 + 0.7)
                + 12.0 * np.cos(2 * np.pi * (x - 0.5 * y) / 5_100.0)
                + 120.0)

    z = elev(x, y) + rng.normal(0.0, 0.5, n)
    pts = np.stack([x, y, z], 1).astype(np.float32)
    idx = rng.integers(0, n, nq)
    q = (pts[idx].astype(np.float64) + rng.normal(0.0, 1.0, (nq, 3))).astype(np.float32)
    return pts, q


def _ray_boxes(o, d, lo, hi):
    """Nearest positive hit distance of rays (o, d) against axis-aligned boxes."""
    inv = 1.0 / np.where(np.abs(d) < 1e-12, 1e-12, d)
    best = np.full(len(d), np.inf)
    for b in range(len(lo)):
        t1 = (lo[b] - o) * inv
        t2 = (hi[b] - o) * inv
        tmin = np.minimum(t1, t2).max(axis=1)
        tmax = np.maximum(t1, t2).min(axis=1)
        hit = (tmax >= np.maximum(tmin, 0.0)) & (tmin > 0.0)
        best = np.where(hit & (tmin < best), tmin, best)
    return best


def lidar_scan(frames=20, rings=64, az=1024, nq=400_000, seed=3, noise=0.02, n_boxes=30):
    rng = np.random.default_rng(seed)
    h = 1.73
    el = np.deg2rad(np.linspace(-24.8, 2.0, rings))
    centers = np.stack([rng.uniform(-50, 70, n_boxes), rng.uniform(-60, 60, n_boxes)], 1)
    size = np.stack([rng.uniform(1.5, 12, n_boxes), rng.uniform(1.5, 12, n_boxes), rng.uniform(1.2, 8, n_boxes)], 1)
    lo = np.concatenate([centers - size[:, :2] / 2, np.zeros((n_boxes, 1))], 1)
    hi = np.concatenate([centers + size[:, :2] / 2, size[:, 2:3]], 1)
    out = []
    for f in range(frames):
        o = np.array([1.0 * f, 0.02 * f, h])
        a = np.linspace(0, 2 * np.pi, az, endpoint=False) + rng.uniform(0, 2 * np.pi / az)
        E, A = np.meshgrid(el, a, indexing="ij")
        d = np.stack([np.cos(E) * np.cos(A), np.cos(E) * np.sin(A), np.sin(E)], -1).reshape(-1, 3)
        t = np.full(len(d), np.inf)
        down = d[:, 2] < 0
        t[down] = h / -d[down, 2]
        # cylinder of radius 120 m around the sensor
        horiz = np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2)
        tcyl = 120.0 / np.maximum(horiz, 1e-9)
        t = np.minimum(t, tcyl)
        tb = _ray_boxes(o[None, :], d, lo, hi)
        t = np.minimum(t, tb)
        t = t + rng.normal(0.0, noise, len(t))
        out.append(o[None, :] + d * t[:, None])
    pts = np.concatenate(out).astype(np.float32)
    idx = rng.integers(0, len(pts), nq)
    q = (pts[idx].astype(np.float64) + rng.normal(0.0, noise, (nq, 3))).astype(np.float32)
    return pts, q


def _mixture(rng, n, centers, sigmas):
    nb = int(round(0.05 * n))
    nc = n - nb
    which = rng.integers(0, len(centers), nc)
    cl = centers[which] + rng.normal(0.0, 1.0, (nc, 3)) * sigmas[which, None]
    bg = rng.random((nb, 3))
    p = np.concatenate([cl, bg])
    rng.shuffle(p, axis=0)
    return p.astype(np.float32)


def clustered_batch(n_clouds=256, lo=1e3, hi=2e5, seed=4, ks=(1, 5, 16, 64)):
    """Returns (pts, cloud_offsets, queries, query_offsets, k_per_cloud)."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), n_clouds))).astype(np.int64)
    P, Q, kc = [], [], []
    for c in range(n_clouds):
        m = int(rng.integers(1, 51))
        centers = rng.random((m, 3))
        sig = np.exp(rng.uniform(np.log(1e-3), np.log(1e-1), m))
        P.append(_mixture(rng, int(sizes[c]), centers, sig))
        Q.append(_mixture(rng, int(round(0.1 * sizes[c])), centers, sig))
        kc.append(int(rng.choice(ks)))
    poff = np.concatenate([[0], np.cumsum([len(p) for p in P])]).astype(np.int64)
    qoff = np.concatenate([[0], np.cumsum([len(q) for q in Q])]).astype(np.int64)
    return np.concatenate(P), poff, np.concatenate(Q), qoff, np.array(kc, np.int32)


def plummer(n=1_000_000, a=1.0, seed=5, rmax=100.0):
    """Plummer sphere, equal masses summing to 1; radius truncated at rmax*a.
    Returns float32 (n, 4) = (x, y, z, m)."""
    rng = np.random.default_rng(seed)
    r = np.empty(n)
    filled = 0
    while filled < n:
        u = rng.random(n - filled)
        u = u[u > 0]
        rr = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        rr = rr[rr <= rmax * a]
        r[filled: filled + len(rr)] = rr
        filled += len(rr)
    cth = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    sth = np.sqrt(1 - cth ** 2)
    pos = np.stack([r * sth * np.cos(phi), r * sth * np.sin(phi), r * cth], 1)
    m = np.full((n, 1), 1.0 / n)
    return np.concatenate([pos, m], 1).astype(np.float32)
