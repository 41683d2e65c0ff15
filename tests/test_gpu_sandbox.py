"""Sandbox harnesses end to end on the GPU (gprt-sandbox/src/main.rs,
src/bin/barneshut.rs): text files in, the reference's output lines and
knn_results.txt out, ids checked against the FP64 oracle."""
import numpy as np
import pytest

import gprt_hip
import oracle
from gprt_hip import barneshut, sandbox

pytestmark = pytest.mark.gpu


def _csv(path, a):
    np.savetxt(path, a, delimiter=",", fmt="%.9g")
    return str(path)


def _read_dump(path):
    rows = []
    for line in open(path):
        head, body = line.rstrip("\n").split(":", 1)
        assert head == f"Q{len(rows)}"
        body = body.strip("[]")
        rows.append([int(x) for x in body.split(", ")] if body else [])
    return rows


def test_sandbox_knn_dump(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    rng = np.random.default_rng(21)
    pts = rng.normal(size=(40_000, 3)).astype(np.float32)
    q = rng.normal(size=(3_000, 3)).astype(np.float32)
    dp, qp = _csv(tmp_path / "data.csv", pts), _csv(tmp_path / "queries.csv", q)
    lines = []
    sandbox.run(["gprt-sandbox", dp, qp, "7", "0.3", "2.0", "dump"], lines.append)
    assert lines[0].startswith("[IO] Loaded 40000 points and 3000 queries in ")
    assert lines[1] == "[LOAD] Using CLI defaults: P=0.30, M=2.0"
    assert any(l.startswith("[GPRT_LOOP] Initial Radius Calculated: ") for l in lines)
    stats = [l for l in lines if l.startswith("[GPRT_STATS] search_ms=")]
    assert len(stats) == 1 and stats[0].endswith("%")
    last = lines[-1].split(",")
    assert last[0] == dp and last[1:3] == ["3000", "7"] and float(last[3]) > 0 and last[4] == str(3000 * 7)
    # the text round trip of the inputs is exact (%.9g), so the oracle sees the same points
    pts_r, q_r = gprt_hip.read_points(dp), gprt_hip.read_points(qp)
    assert np.array_equal(pts_r, pts) and np.array_equal(q_r, q)
    ref = oracle.knn_brute(pts, q, 7)
    assert _read_dump(tmp_path / "knn_results.txt") == ref.tolist()


def test_sandbox_small_cloud_and_wisdom(tmp_path, monkeypatch):
    """Fewer points than k: short lists; a reference-form wisdom file is loaded."""
    monkeypatch.chdir(tmp_path)
    rng = np.random.default_rng(22)
    pts = rng.random((3, 3), dtype=np.float32)
    q = rng.random((50, 3), dtype=np.float32)
    dp, qp = _csv(tmp_path / "d.csv", pts), _csv(tmp_path / "q.csv", q)
    (tmp_path / "gprt_wisdom.json").write_text('{"radius_increment_mult": 2.0, "max_hits_per_query": 2000, '
                                               '"use_morton_lbv": true, "radius_heuristic": '
                                               '{"SampledPercentile": 0.2}, "memory_strategy": '
                                               '"PayloadRegisterHeap"}')
    lines = []
    sandbox.run(["s", dp, qp, "5", "dump"], lines.append)
    assert lines[1] == "[LOAD] Loaded auto-tuned schedule from gprt_wisdom.json"
    assert lines[-1].endswith(f",50,5,{lines[-1].split(',')[3]},150")
    ref = oracle.knn_brute(pts, q, 5)
    assert _read_dump(tmp_path / "knn_results.txt") == [r[:3] for r in ref.tolist()]


def test_sandbox_tune_then_load(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    rng = np.random.default_rng(23)
    pts = rng.random((30_000, 3), dtype=np.float32)
    q = rng.random((5_000, 3), dtype=np.float32)
    dp, qp = _csv(tmp_path / "d.csv", pts), _csv(tmp_path / "q.csv", q)
    lines = []
    ids = sandbox.run(["s", dp, qp, "5", "tune"], lines.append)
    assert "[TUNE] Saved optimal schedule to gprt_wisdom.json" in lines
    assert (tmp_path / "gprt_wisdom.json").exists()
    ref = oracle.knn_brute(pts, q, 5)
    assert np.array_equal(ids, ref)
    lines = []
    ids = sandbox.run(["s", dp, qp, "5"], lines.append)
    assert lines[1] == "[LOAD] Loaded auto-tuned schedule from gprt_wisdom.json"
    assert np.array_equal(ids, ref)


def test_barneshut_sandbox(tmp_path):
    rng = np.random.default_rng(24)
    n = 4_000
    pm = np.concatenate([rng.normal(size=(n, 3)), rng.random((n, 1)) + 0.5], axis=1).astype(np.float32)
    # a pair 1e-3 apart (where the reference's law is twice the Plummer one)
    # and a coincident duplicate (a pair the reference's law skips)
    pm[1, :3] = pm[0, :3] + np.float32([1e-3, 0, 0])
    pm[2] = pm[3]
    path = tmp_path / "galaxy.txt"
    with open(path, "w") as f:
        f.write("0,0,0,0\n")  # the dummy line the reader skips
        np.savetxt(f, pm, delimiter=",", fmt="%.9g")
    lines = []
    forces = barneshut.run(["barneshut", str(path), "0.0"], lines.append)
    assert lines[:3] == [f"Loading Galaxy dataset from {path}...", f"Loaded {n} bodies.",
                         "Compiling and launching RT-BarnesHut via GPRT DSL..."]
    assert lines[3].startswith("Total Force Magnitude: ")
    assert f"Forces Computed: {n}" in lines
    # theta = 0 visits every body: the FP64 direct sum of the reference's law
    acc = oracle.bh_direct(pm, G=barneshut.G_NEWTON, eps=barneshut.EPS, law="reference")
    want = acc * pm[:, 3:4].astype(np.float64)
    err = np.linalg.norm(forces - want, axis=1) / np.linalg.norm(want, axis=1)
    assert err.max() < 2e-5
    plummer = oracle.bh_direct(pm[:2], G=barneshut.G_NEWTON, eps=barneshut.EPS) * pm[:2, 3:4]
    assert np.linalg.norm(forces[:2] - plummer, axis=1).min() > 0.1 * np.linalg.norm(plummer, axis=1).min()
    total = np.linalg.norm(want, axis=1).sum()
    got = float(lines[3].split(": ")[1])
    assert abs(got - total) / total < 1e-3


def test_barneshut_reference_physics_check(tmp_path):
    """The reference's own Barnes-Hut check (gprt-sandbox/verify_physics.py:5-48):
    10,000 bodies from RandomState(42), masses U(0.5, 1.5) drawn first, then
    positions U(-100, 100)^3, written as "x,y,z,m" lines; `barneshut <file> 0.5`
    must print a total |F| within 2% of the FP64 O(N^2) Newtonian sum (pairs
    closer than 1e-6 skipped).  Per-body errors are checked here as well."""
    n = 10_000
    rs = np.random.RandomState(42)
    mass = rs.uniform(0.5, 1.5, n)
    pos = rs.uniform(-100, 100, (n, 3))
    path = tmp_path / "small_test.txt"
    path.write_text("".join(f"{x},{y},{z},{m}\n" for (x, y, z), m in zip(pos.tolist(), mass.tolist())))
    lines = []
    forces = barneshut.run(["barneshut", str(path), "0.5"], lines.append)
    # FP64 ground truth from the same (float64) values, 250 bodies at a time
    exact = np.empty((n, 3))
    for i0 in range(0, n, 250):
        d = pos[None, :, :] - pos[i0:i0 + 250, None, :]
        r = np.sqrt((d * d).sum(-1))
        w = np.where(r > 1e-6, barneshut.G_NEWTON * mass[i0:i0 + 250, None] * mass[None, :] /
                     np.maximum(r, 1e-6) ** 3, 0.0)
        exact[i0:i0 + 250] = (d * w[..., None]).sum(1)
    mag_exact = np.linalg.norm(exact, axis=1).sum()
    total = float(lines[3].split(": ")[1])
    assert abs(total - mag_exact) / mag_exact < 0.02
    err = np.linalg.norm(forces - exact, axis=1) / np.linalg.norm(exact, axis=1)
    print(f"total |F| rel. error {abs(total - mag_exact) / mag_exact:.2e}; per body median {np.median(err):.2e} "
          f"p99 {np.quantile(err, 0.99):.2e}")
    assert np.median(err) < 5e-3 and np.quantile(err, 0.99) < 5e-2
