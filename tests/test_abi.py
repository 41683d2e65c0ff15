"""CPU-side checks of the C ABI: the library loads, exports every entry point
that include/gprt_hip.h declares, and rejects bad arguments on the host
before touching a device.  No compute runs here (no GPU in this container)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import gprt_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gprt_hip.h")


@pytest.mark.parametrize("lang,std", [("c", "-std=c99"), ("c++", "-std=c++11")])
def test_header_is_plain_c_and_cxx(lang, std, tmp_path):
    """The boundary header compiles on its own as C99 (what cgo, a C binding
    generator or ctypes users read) and as C++11, without warnings."""
    import shutil
    cc = shutil.which("gcc" if lang == "c" else "g++")
    if cc is None:
        pytest.skip("no host compiler")
    r = subprocess.run([cc, "-fsyntax-only", std, "-Wall", "-Wextra", "-Werror", "-x", lang, HEADER],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def declared_symbols():
    text = open(HEADER).read()
    return re.findall(r"GPRT_HIP_API\s+[\w\s\*]+?\b(\w+)\s*\(", text)


def test_header_declares_the_reference_ffi():
    names = set(declared_symbols())
    # gprt-optix/src/lib.rs:13-31
    ref = {"gprt_pipeline_create", "gprt_pipeline_destroy", "gprt_bvh_build", "gprt_bvh_build_triangles",
           "gprt_bvh_refit", "gprt_bvh_destroy", "gprt_register_array", "gprt_register_value", "gprt_execute",
           "gprt_retrieve_array_lengths", "gprt_retrieve_array_flat", "gprt_retrieve_value",
           "gprt_execute_autorope_soa_ffi"}
    assert ref <= names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(gprt_hip.LIB_PATH), "build the library first (python rt-dsl_amd/build.py)"
    out = subprocess.run(["nm", "-D", "--defined-only", gprt_hip.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    names = {n for n, _, _ in gprt_hip.SIGNATURES}
    assert set(declared_symbols()) <= names


def test_host_side_entry_points_without_gpu():
    L = gprt_hip.lib()
    assert L.gprt_hip_version() >= 10000
    cfg = gprt_hip.default_config()
    # defaults mirror the documented Schedule superset
    assert cfg.sampler == gprt_hip.SAMPLER_LCG
    assert cfg.radius_heuristic == gprt_hip.RADIUS_PERCENTILE
    assert abs(cfg.radius_param - 0.95) < 1e-7 and cfg.radius_mult == 3.0
    assert cfg.sort_queries == 1 and cfg.use_graph == 1
    n = ctypes.c_int(-1)
    assert L.gprt_hip_num_devices(ctypes.byref(n)) == 0
    assert n.value >= 0


def test_null_arguments_are_rejected_before_device_work():
    L = gprt_hip.lib()
    assert L.gprt_hip_ctx_create(0, None) == gprt_hip.ERR_ARG
    assert L.gprt_hip_knn(None, None, None, 0, 5, None, None, None, 0) == gprt_hip.ERR_ARG
    assert L.gprt_hip_knn_plain(None, None, None, 0, 5, None, None, 0) == gprt_hip.ERR_ARG
    assert L.gprt_hip_range(None, None, None, 0, 1.0, None, None, 0, 0) == gprt_hip.ERR_ARG
    assert L.gprt_hip_bh_forces(None, None, 0, 0.5, 1e-3, 1.0, 8, None, 0) == gprt_hip.ERR_ARG
    assert L.gprt_hip_sort_pairs(None, None, None, 0, 0, 64, 0) == gprt_hip.ERR_ARG
    assert b"null" in L.gprt_hip_last_error()
    assert L.gprt_hip_ctx_destroy(None) == 0
    assert L.gprt_hip_tree_destroy(None) == 0


def test_context_creation_fails_cleanly_without_gpu():
    n = ctypes.c_int(0)
    gprt_hip.lib().gprt_hip_num_devices(ctypes.byref(n))
    if n.value > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(gprt_hip.GprtError) as e:
        gprt_hip.Context(0)
    assert e.value.code == gprt_hip.ERR_HIP


def test_wrapper_validates_shapes():
    with pytest.raises(ValueError):
        gprt_hip._points(np.zeros((5, 2), np.float32))
    assert gprt_hip._points(np.zeros((5, 3), np.float64)).dtype == np.float32
