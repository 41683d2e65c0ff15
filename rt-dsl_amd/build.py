"""Build libgprt_hip.so (all HIP translation units, gfx950 only) in-tree.

    python rt-dsl_amd/build.py [-j N] [--incremental]

Objects go to rt-dsl_amd/_build/, the library to rt-dsl_amd/gprt_hip/, together
with libgprt_host.so (the k-NN walk built for the host, for CPU tests).
--sanitize builds both with host-side ASan + UBSan into rt-dsl_amd/_build_asan/.
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

SOURCES = ["lbvh.hip", "sah.hip", "knn.hip", "knn_lane.hip", "knn_wide.hip", "knn_packet.hip", "knn_qwave.hip", "api.hip", "range.hip",
           "bh.hip", "legacy.hip", "autotune.hip", "codegen.hip", "io.hip"]


HOST_OUT = os.path.join(HERE, "gprt_hip", "libgprt_host.so")
# host-side AddressSanitizer + UndefinedBehaviorSanitizer (GPU code is built as
# usual; every -fsanitize applies to the host pass only)
SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined", "-g"]
SAN_DIR = os.path.join(HERE, "_build_asan")


def compile_one(src, build_dir=BUILD, defines=(), extra=()):
    t0 = time.time()
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + HIPFLAGS + list(extra) + ["-D" + d for d in defines] + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    return obj, err + f"\n[{src}: {time.time() - t0:.0f} s]"


def _up_to_date(src, build_dir):
    """True when the object of `src` is newer than it and than every header
    (csrc/*.hpp and include/gprt_hip.h)."""
    obj = os.path.join(build_dir, os.path.splitext(src)[0] + ".o")
    if not os.path.exists(obj):
        return False
    deps = [os.path.join(CSRC, src), os.path.join(HERE, "..", "include", "gprt_hip.h")]
    deps += [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".hpp")]
    return os.path.getmtime(obj) > max(os.path.getmtime(d) for d in deps)


def build(jobs=None, verbose=True, out=OUT, build_dir=BUILD, defines=(), sanitize=False, incremental=False):
    """Compile every translation unit and link the library.  `defines` are
    extra -D options (tuning constants of the kernels), for builds of the
    same library into another `out` / `build_dir`.  sanitize=True builds the
    host code with ASan + UBSan (load it with the shared ASan runtime
    preloaded).  incremental=True skips translation units whose object is
    newer than their source and every header (same defines assumed)."""
    os.makedirs(build_dir, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 8)
    extra = SANITIZE if sanitize else []

    def one(src):
        if incremental and _up_to_date(src, build_dir):
            return os.path.join(build_dir, os.path.splitext(src)[0] + ".o"), ""
        return compile_one(src, build_dir, defines, extra)

    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(one, srcs))
    for obj, err in results:
        if err.strip() and verbose:
            print(err, file=sys.stderr)
    tmp = out + ".tmp"
    link_san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-shared-libsan"] \
        if sanitize else []
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + link_san + [o for o, _ in results] + \
        ["-lhiprtc"]
    subprocess.check_call(cmd)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_host(out=HOST_OUT, sanitize=False):
    """libgprt_host.so: the per-lane k-NN walk and the LBVH layout compiled for
    the host only (host_walk.cpp), for CPU tests of the device code's logic."""
    cmd = [HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
           "-fvisibility=hidden", "-Wall"]
    if sanitize:
        cmd += ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-Xarch_host",
                "-fno-sanitize-recover=undefined", "-shared-libsan", "-g"]
    tmp = out + ".tmp"
    r = subprocess.run(cmd + [os.path.join(CSRC, "host_walk.cpp"), "-o", tmp], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host build failed:\n" + r.stderr)
    os.replace(tmp, out)
    return out


def asan_runtime():
    """Path of the shared ASan runtime of the HIP toolchain (to LD_PRELOAD)."""
    clang = os.path.join(os.path.dirname(os.path.realpath(HIPCC)), "..", "lib", "llvm", "bin", "clang++")
    if not os.path.exists(clang):
        clang = "/opt/rocm/llvm/bin/clang++"
    return subprocess.run([clang, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True,
                          text=True).stdout.strip()


def build_sanitized(jobs=None):
    """The library and the host walk with ASan + UBSan on the host side, into
    rt-dsl_amd/_build_asan/."""
    os.makedirs(SAN_DIR, exist_ok=True)
    lib = build(jobs, verbose=False, out=os.path.join(SAN_DIR, "libgprt_hip.so"), build_dir=SAN_DIR, sanitize=True)
    host = build_host(out=os.path.join(SAN_DIR, "libgprt_host.so"), sanitize=True)
    return lib, host


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-D", dest="defines", action="append", default=[], help="extra preprocessor define")
    ap.add_argument("--out", default=OUT)
    ap.add_argument("--build-dir", default=BUILD)
    ap.add_argument("--sanitize", action="store_true", help="ASan + UBSan host build into _build_asan/")
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only sources changed since their object (or all, after a header change)")
    a = ap.parse_args()
    if a.sanitize:
        print(build_sanitized(a.j))
    else:
        build(a.j, out=a.out, build_dir=a.build_dir, defines=a.defines, incremental=a.incremental)
        build_host()
