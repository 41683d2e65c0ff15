"""The CPU suite again with the host code of both libraries built under
AddressSanitizer + UndefinedBehaviorSanitizer (python rt-dsl_amd/build.py
--sanitize): the JSON and point readers, the code generator, the ABI checks
and the host build of the k-NN walk run instrumented, in a child process with
the toolchain's ASan runtime preloaded.  Skipped when that build is absent."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "rt-dsl_amd", "_build_asan")


def _sanitized_build_state():
    libs = [os.path.join(SAN, "libgprt_hip.so"), os.path.join(SAN, "libgprt_host.so")]
    if not all(os.path.exists(p) for p in libs):
        return "no sanitizer build (python rt-dsl_amd/build.py --sanitize)"
    srcs = [os.path.join(d, f) for d in (os.path.join(ROOT, "rt-dsl_amd", "csrc"), os.path.join(ROOT, "include"))
            for f in os.listdir(d)]
    if max(os.path.getmtime(p) for p in srcs) > min(os.path.getmtime(p) for p in libs):
        return "the sanitizer build is older than the sources (python rt-dsl_amd/build.py --sanitize)"
    return None


@pytest.mark.skipif(_sanitized_build_state() is not None, reason=str(_sanitized_build_state()))
def test_cpu_suite_under_asan_ubsan():
    sys.path.insert(0, os.path.join(ROOT, "rt-dsl_amd"))
    import build
    env = dict(os.environ, LD_PRELOAD=build.asan_runtime(), ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1", GPRT_HIP_LIBRARY=os.path.join(SAN, "libgprt_hip.so"),
               GPRT_HOST_LIBRARY=os.path.join(SAN, "libgprt_host.so"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                        "--ignore", os.path.abspath(__file__), os.path.join(ROOT, "tests")],
                       env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail
