"""ctypes front end of libgprt_host.so: the per-lane FP32 k-NN walk, its
candidate list and FP64 re-rank (lane_walk.hpp), the Barnes-Hut walk and mass
refit (bh_walk.hpp, bh.hip) and the LBVH layout of lbvh.hip, compiled for the
host with every memory access bounds-checked.

For tests of the device code's logic on the CPU; nothing in the GPU path uses
it.  GPRT_HOST_LIBRARY selects another build (e.g. the ASan + UBSan one).
"""
import ctypes
import os

import numpy as np

_PATH = os.environ.get("GPRT_HOST_LIBRARY",
                       os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgprt_host.so"))
_lib = None
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

CAPACITIES = (8, 16, 24, 32, 40, 48, 64, 67)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise RuntimeError(f"{_PATH} is not built (python rt-dsl_amd/build.py)")
        L = ctypes.CDLL(_PATH)
        L.gprt_host_last_error.restype = ctypes.c_char_p
        L.gprt_host_build_tree.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, _u64p, _u32p, _f32p, _i32p]
        L.gprt_host_lane_knn.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, _f32p, ctypes.c_int64, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                         _u32p, _i32p, _i64p]
        L.gprt_host_build_tree_ex.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _u64p, _u32p,
                                              _f32p, _i32p, _i32p, _i32p]
        L.gprt_host_lane_knn_ex.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _f32p,
                                            ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                            ctypes.c_float, ctypes.c_int, ctypes.c_int, _u32p, _i32p, _i64p]
        L.gprt_host_bh.argtypes = [_f32p, ctypes.c_int64, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, _f32p, _f64p,
                                   _f64p, _f64p, _i64p]
        _lib = L
    return _lib


def build_tree(pts, leaf_size=8):
    """(keys, perm, nodes (n-1, 16) float32, root (2,)) as lbvh.hip lays them out."""
    pts = np.ascontiguousarray(pts, np.float32)
    n = len(pts)
    keys = np.zeros(n, np.uint64)
    perm = np.zeros(n, np.uint32)
    nodes = np.zeros((max(n - 1, 1), 16), np.float32)
    root = np.zeros(2, np.int32)
    rc = lib().gprt_host_build_tree(pts, n, leaf_size, keys, perm, nodes, root)
    assert rc == 0
    return keys, perm, nodes[: max(n - 1, 0)], root


def build_tree_ex(pts, leaf_size=8, sah=False):
    """(keys, perm, nodes (n-1, 16), root (2,), parent (2n-1,), ranges (n-1, 2)) of
    the LBVH or (sah=True) of the SAH tree of sah.hip, as the GPU lays them out;
    ranges of unused node ids are [0, -1]."""
    pts = np.ascontiguousarray(pts, np.float32)
    n = len(pts)
    keys = np.zeros(n, np.uint64)
    perm = np.zeros(n, np.uint32)
    nodes = np.zeros((max(n - 1, 1), 16), np.float32)
    root = np.zeros(2, np.int32)
    parent = np.full(max(2 * n - 1, 1), -1, np.int32)
    ranges = np.zeros((max(n - 1, 1), 2), np.int32)
    rc = lib().gprt_host_build_tree_ex(pts, n, leaf_size, 1 if sah else 0, keys, perm, nodes, root, parent, ranges)
    if rc:
        raise RuntimeError(lib().gprt_host_last_error().decode())
    m = max(n - 1, 0)
    return keys, perm, nodes[:m], root, parent[: max(2 * n - 1, 0)], ranges[:m]


def lane_knn(pts, queries, k, cap, leaf_size=8, radius0=1.0, mult=3.0, rounds=3, stack_depth=64, sah=False):
    """The batched form's per-lane path, one query at a time.  Returns (ids
    (nq, k), status (nq,): 0 written by the re-rank, 1 sent to the exact kernel,
    2 id guard tripped, stats: node visits, point tests, trips).  Raises on an
    out-of-bounds access."""
    pts = np.ascontiguousarray(pts, np.float32)
    q = np.ascontiguousarray(queries, np.float32)
    ids = np.full((len(q), k), 0xFFFFFFFF, np.uint32)
    st = np.zeros(len(q), np.int32)
    stats = np.zeros(3, np.int64)
    rc = lib().gprt_host_lane_knn_ex(pts, len(pts), leaf_size, 1 if sah else 0, q, len(q), k, cap, radius0, mult,
                                     rounds, stack_depth, ids, st, stats)
    if rc:
        raise RuntimeError(lib().gprt_host_last_error().decode())
    return ids, st, stats


def barnes_hut(pos_mass, theta, eps=1e-3, G=1.0, leaf_size=8, law="plummer", sample=None):
    """Barnes-Hut of the bodies `sample` (original indices, default all) as the
    GPU kernel computes it.  Returns a dict: acc32 (ns, 3) float32, the kernel's
    FP32 operations; acc64 (ns, 3), the same accepted children and bodies
    summed in FP64; bound (ns,), a derived bound on |acc64 - direct sum| per
    body (NaN for the reference law); absum (ns,), the sum of the magnitudes
    of the summed terms; stats (visits, bodies, accepted)."""
    pm = np.ascontiguousarray(pos_mass, np.float32)
    sel = None if sample is None else np.ascontiguousarray(sample, np.int64)
    ns = len(pm) if sel is None else len(sel)
    acc32 = np.zeros((ns, 3), np.float32)
    acc64 = np.zeros((ns, 3), np.float64)
    bound = np.zeros(ns, np.float64)
    absum = np.zeros(ns, np.float64)
    stats = np.zeros(3, np.int64)
    rc = lib().gprt_host_bh(pm, len(pm), leaf_size, theta, eps, G, 1 if law == "reference" else 0,
                            None if sel is None else sel.ctypes.data, ns, acc32, acc64, bound, absum, stats)
    if rc:
        raise RuntimeError(lib().gprt_host_last_error().decode())
    return {"acc32": acc32, "acc64": acc64, "bound": bound, "absum": absum, "stats": stats}
