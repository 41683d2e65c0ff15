"""Build libgprt_hip.so (all HIP translation units, gfx950 only) in-tree.

    python rt-dsl_amd/build.py [-j N]

Objects go to rt-dsl_amd/_build/, the library to rt-dsl_amd/gprt_hip/.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "gprt_hip", "libgprt_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: products and sums are rounded separately unless an fma is
# written explicitly (the distance contract depends on it).
# -packed-fp32-ops off: no v_pk_*_f32 generation.
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]

SOURCES = ["lbvh.hip", "knn.hip", "knn_lane.hip", "knn_wide.hip", "knn_packet.hip", "knn_qwave.hip", "api.hip", "range.hip",
           "bh.hip", "legacy.hip", "autotune.hip", "codegen.hip", "io.hip"]


def compile_one(src, build_dir=BUILD, defines=()):
    t0 = time.time()
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + HIPFLAGS + ["-D" + d for d in defines] + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    return obj, err + f"\n[{src}: {time.time() - t0:.0f} s]"


def build(jobs=None, verbose=True, out=OUT, build_dir=BUILD, defines=()):
    """Compile every translation unit and link the library.  `defines` are
    extra -D options (tuning constants of the kernels), for builds of the
    same library into another `out` / `build_dir`."""
    os.makedirs(build_dir, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda src: compile_one(src, build_dir, defines), srcs))
    for obj, err in results:
        if err.strip() and verbose:
            print(err, file=sys.stderr)
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [o for o, _ in results] + ["-lhiprtc"]
    subprocess.check_call(cmd)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-D", dest="defines", action="append", default=[], help="extra preprocessor define")
    ap.add_argument("--out", default=OUT)
    ap.add_argument("--build-dir", default=BUILD)
    a = ap.parse_args()
    build(a.j, out=a.out, build_dir=a.build_dir, defines=a.defines)
