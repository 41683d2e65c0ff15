"""Barnes-Hut command-line harness with the reference sandbox's arguments and output.

    python -m gprt_hip.barneshut <dataset.txt> <theta>

Follows gprt-sandbox/src/bin/barneshut.rs:36-70: reads "x,y,z,m" lines
(gprt_hip.read_points with CSV_BODIES, the rules of load_galaxy,
barneshut.rs:8-34), times the force computation (tree build, walk, copies in
and out), and prints the sum over bodies of |F_i| and a report block.  F_i is
m_i times the library's acceleration under the reference's force law
(gprt-macros/src/lib.rs:785, 790, 799): G m_i m_j (x_j - x_i) /
(d^2 sqrt(d^2 + 1e-6)) over pairs with d^2 > 1e-10, G = 6.674e-11, summed
with the Barnes-Hut approximation at the given theta (theta = 0 sums every
pair, which is what the reference's launch does as shipped).
"""
import math
import sys
import time

import numpy as np

import gprt_hip

G_NEWTON = 6.674e-11
EPS = 1e-3


def rust_exp(v, digits=4):
    """Rust's "{:.4e}" form: 1.2346e-3, 5.0000e0, -2.5000e11."""
    if v == 0.0:
        return f"{0.0:.{digits}f}e0"
    if not math.isfinite(v):
        return "NaN" if math.isnan(v) else ("inf" if v > 0 else "-inf")
    m, e = f"{v:.{digits}e}".split("e")
    return f"{m}e{int(e)}"


def run(argv, out=print):
    if len(argv) < 3:
        print(f"Usage: {argv[0] if argv else 'barneshut'} <dataset.txt> <theta>", file=sys.stderr)
        raise SystemExit(1)
    path = argv[1]
    try:
        theta = float(argv[2])
    except ValueError:
        raise SystemExit("Theta must be a float")
    out(f"Loading Galaxy dataset from {path}...")
    bodies = gprt_hip.read_points(path, gprt_hip.CSV_BODIES)
    out(f"Loaded {len(bodies)} bodies.")
    out("Compiling and launching RT-BarnesHut via GPRT DSL...")
    ctx = gprt_hip.Context(0)
    try:
        t0 = time.perf_counter()
        acc = (ctx.barnes_hut(bodies, theta=theta, eps=EPS, G=G_NEWTON, law="reference") if len(bodies)
               else np.zeros((0, 3), np.float32))
        ms = (time.perf_counter() - t0) * 1e3
    finally:
        ctx.close()
    forces = (acc * bodies[:, 3:4]).astype(np.float32)
    mag = np.sqrt((forces * forces).sum(axis=1, dtype=np.float32))
    total = float(np.float32(mag.sum(dtype=np.float32)))
    out(f"Total Force Magnitude: {rust_exp(total)}")
    out("\n=== GPRT DSL BARNES-HUT REPORT ===")
    out(f"Execution Time : {ms:.3f} ms")
    out(f"Forces Computed: {len(forces)}")
    out("==================================\n")
    return forces


def main(argv=None):
    run(sys.argv if argv is None else argv)


if __name__ == "__main__":
    main()
