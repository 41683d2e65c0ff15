"""ctypes front end of the CPU oracle (oracle.c).

Test infrastructure only: tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py import this module as the checker.  The product library never
imports it.  Build the shared object with `make -C oracle` (done by
__graft_entry__.build()).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64
_int = ctypes.c_int


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_knn_brute.argtypes = [_f32p, _i64, _f32p, _i64, _int, _u32p, ctypes.c_void_p, _int]
        L.oracle_range_brute.argtypes = [_f32p, _i64, _f32p, _i64, ctypes.c_float, _i64p, ctypes.c_void_p, _int]
        L.oracle_expand_bits.argtypes = [ctypes.c_uint32]
        L.oracle_expand_bits.restype = ctypes.c_uint32
        L.oracle_morton3.argtypes = [ctypes.c_uint32] * 3
        L.oracle_morton3.restype = ctypes.c_uint32
        L.oracle_morton_codes.argtypes = [_f32p, _i64, _u32p, ctypes.c_void_p]
        L.oracle_morton_codes_bbox.argtypes = [_f32p, _i64, _f32p, _u32p]
        L.oracle_morton_codes_wide.argtypes = [_f32p, _i64, _f32p, _int, _u64p]
        L.oracle_lcg_draws.argtypes = [ctypes.c_uint64, _i64, _u64p]
        L.oracle_lcg_draws.restype = ctypes.c_uint64
        L.oracle_sampler_lcg.argtypes = [_f32p, _i64, _f32p, _i64, _int, ctypes.c_float]
        L.oracle_sampler_lcg.restype = ctypes.c_float
        L.oracle_sampler_strided.argtypes = [_f32p, _i64, _f32p, _i64, _int, _int, ctypes.c_float]
        L.oracle_sampler_strided.restype = ctypes.c_float
        L.oracle_bh_direct.argtypes = [_f32p, _i64, ctypes.c_void_p, _i64, ctypes.c_double, ctypes.c_double,
                                       _f64p, _int]
        L.oracle_bh_direct_ref.argtypes = L.oracle_bh_direct.argtypes
        L.oracle_kdtree_knn.argtypes = [_f32p, _i64, _f32p, _i64, _int, _u32p, ctypes.c_void_p, _int,
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.oracle_num_threads.restype = _int
        L.oracle_d2_64.argtypes = [_f32p, _f32p]
        L.oracle_d2_64.restype = ctypes.c_double
        _lib = L
    return _lib


def _pts(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError("expected an (n, 3) float32 array")
    return a


def knn_brute(pts, qs, k, with_d2=False, nthreads=0):
    """Exact k-NN, ids ordered by (d64, id); pads with 0xFFFFFFFF when n < k."""
    pts, qs = _pts(pts), _pts(qs)
    ids = np.empty((len(qs), k), np.uint32)
    d2 = np.empty((len(qs), k), np.float64) if with_d2 else None
    rc = lib().oracle_knn_brute(pts, len(pts), qs, len(qs), k, ids,
                                d2.ctypes.data if with_d2 else None, nthreads)
    assert rc == 0
    return (ids, d2) if with_d2 else ids


def range_brute(pts, qs, r, nthreads=0):
    """Fixed-radius neighbours (d64 <= r*r in binary64); CSR (offsets, ids)."""
    pts, qs = _pts(pts), _pts(qs)
    off = np.zeros(len(qs) + 1, np.int64)
    L = lib()
    L.oracle_range_brute(pts, len(pts), qs, len(qs), float(r), off, None, nthreads)
    ids = np.empty(int(off[-1]), np.uint32)
    if len(ids):
        L.oracle_range_brute(pts, len(pts), qs, len(qs), float(r), off, ids.ctypes.data, nthreads)
    return off, ids


def expand_bits(x):
    return lib().oracle_expand_bits(x)


def morton3(x, y, z):
    return lib().oracle_morton3(x, y, z)


def morton_codes(pts):
    pts = _pts(pts)
    codes = np.empty(len(pts), np.uint32)
    bbox = np.empty(6, np.float32)
    lib().oracle_morton_codes(pts, len(pts), codes, bbox.ctypes.data)
    return codes, bbox


def morton_codes_bbox(pts, bbox6):
    pts = _pts(pts)
    codes = np.empty(len(pts), np.uint32)
    lib().oracle_morton_codes_bbox(pts, len(pts), np.ascontiguousarray(bbox6, np.float32), codes)
    return codes


def morton_codes_wide(pts, bbox6, bits):
    """Tree sort keys with `bits` per axis (bits = 10: the reference's codes)."""
    pts = _pts(pts)
    codes = np.empty(len(pts), np.uint64)
    if lib().oracle_morton_codes_wide(pts, len(pts), np.ascontiguousarray(bbox6, np.float32), bits, codes):
        raise ValueError("bits must be in [1, 21]")
    return codes


def lcg_draws(seed, count):
    out = np.empty(count, np.uint64)
    lib().oracle_lcg_draws(seed, count, out)
    return out


HEUR_MAX, HEUR_PERCENTILE, HEUR_FIXED = 0, 1, 2


def sampler_lcg(pts, qs, heuristic=HEUR_PERCENTILE, param=0.95):
    pts, qs = _pts(pts), _pts(qs)
    return float(np.float32(lib().oracle_sampler_lcg(pts, len(pts), qs, len(qs), heuristic, param)))


def sampler_strided(pts, qs, k, heuristic=HEUR_PERCENTILE, param=0.10):
    pts, qs = _pts(pts), _pts(qs)
    return float(np.float32(lib().oracle_sampler_strided(pts, len(pts), qs, len(qs), k, heuristic, param)))


def bh_direct(pos_mass, sample=None, G=1.0, eps=1e-3, nthreads=0, law="plummer"):
    """FP64 direct-sum accelerations for the sampled bodies (all when None).
    law "plummer": G m_j d / (|d|^2 + eps^2)^1.5; law "reference": the
    reference shader's G m_j d / (|d|^2 sqrt(|d|^2 + eps^2)) over pairs with
    |d|^2 > 1e-10 (gprt-macros/src/lib.rs:785, 790, 799)."""
    if law not in ("plummer", "reference"):
        raise ValueError("law is 'plummer' or 'reference'")
    pm = np.ascontiguousarray(pos_mass, np.float32)
    assert pm.ndim == 2 and pm.shape[1] == 4
    if sample is None:
        ns, sp = len(pm), None
    else:
        sample = np.ascontiguousarray(sample, np.int64)
        ns, sp = len(sample), sample.ctypes.data
    acc = np.empty((ns, 3), np.float64)
    fn = lib().oracle_bh_direct if law == "plummer" else lib().oracle_bh_direct_ref
    fn(pm, len(pm), sp, ns, G, eps, acc, nthreads)
    return acc


def kdtree_knn(pts, qs, k, nthreads=0, with_d2=False):
    """CPU exact k-NN path; returns (ids, d2 or None, build_ms, query_ms)."""
    pts, qs = _pts(pts), _pts(qs)
    ids = np.empty((len(qs), k), np.uint32)
    d2 = np.empty((len(qs), k), np.float64) if with_d2 else None
    b, q = ctypes.c_double(), ctypes.c_double()
    rc = lib().oracle_kdtree_knn(pts, len(pts), qs, len(qs), k, ids, d2.ctypes.data if with_d2 else None,
                                 nthreads, ctypes.byref(b), ctypes.byref(q))
    assert rc == 0
    return ids, d2, b.value, q.value


def num_threads():
    return lib().oracle_num_threads()
