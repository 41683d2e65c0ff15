"""Python front end of libgprt_hip.so (ctypes over the C ABI in include/gprt_hip.h).

The names follow the reference's DSL (gprt-macros/src/lib.rs): `k_nn`
(k_nn!, :408-604), `range_query` (r_nn!, :729-750), `barnes_hut`
(barnes_hut!, :752-834), `autotune` (gprt_autotune!, :120-214), with the
`Schedule` knobs of gprt-ir/src/lib.rs:58-89 carried by `Config`.

Arrays may be numpy arrays (host; the call copies through) or DeviceArray
buffers resident in HBM (device pointers, GPRT_PTR_DEVICE).  There is no CPU fallback: if
the shared library or the GPU is missing, the calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgprt_hip.so")
# another build of the same library (for A/B measurements); the default is the in-tree build
LIB_PATH = os.environ.get("GPRT_HIP_LIBRARY", LIB_PATH)

OK = 0
ERR_ARG, ERR_INPUT, ERR_HIP, ERR_ALLOC, ERR_CAPACITY, ERR_STATE = 1, 2, 3, 4, 5, 6
PTR_HOST, PTR_DEVICE, NO_SYNC, COLLECT_COUNTERS, KEYS_REF30, BH_REFERENCE_LAW = 0, 1, 2, 4, 8, 16
BUILD_SAH = 64
BH_LANE_WALK = 128  # Barnes-Hut: one independent walk per lane (default: one walk per wavefront)
TIME_KERNELS = 32
ID_NONE = 0xFFFFFFFF
K_MAX = 4096  # GPRT_HIP_K_MAX: largest k of the k-NN entry points
RADIUS_MAX, RADIUS_PERCENTILE, RADIUS_FIXED, RADIUS_INFINITE = 0, 1, 2, 3
SAMPLER_LCG, SAMPLER_STRIDED = 0, 1
TRAV_THREAD, TRAV_WAVE, TRAV_WIDE = 0, 1, 2


class GprtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gprt_hip error {code}: {msg}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [
        ("sampler", ctypes.c_int),
        ("radius_heuristic", ctypes.c_int),
        ("radius_param", ctypes.c_float),
        ("radius_mult", ctypes.c_float),
        ("max_rounds", ctypes.c_int),
        ("traversal", ctypes.c_int),
        ("stack_depth", ctypes.c_int),
        ("guard_slots", ctypes.c_int),
        ("sort_queries", ctypes.c_int),
        ("use_graph", ctypes.c_int),
        ("grid_blocks", ctypes.c_int),
        ("block_threads", ctypes.c_int),
        ("collect_counters", ctypes.c_int),
        ("time_kernels", ctypes.c_int),
        ("lds_stack", ctypes.c_int),
        ("compact_rounds", ctypes.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class RunStats(ctypes.Structure):
    _fields_ = [
        ("rounds_run", ctypes.c_int),
        ("queries", ctypes.c_int64),
        ("fallback_queries", ctypes.c_int64),
        ("partial_queries", ctypes.c_int64),
        ("active_per_round", ctypes.c_int64 * 16),
        ("initial_radius", ctypes.c_float),
        ("ms_total", ctypes.c_double),
        ("ms_query", ctypes.c_double),
        ("ms_build", ctypes.c_double),
        ("node_visits", ctypes.c_int64),
        ("point_tests", ctypes.c_int64),
        ("rerank_reads", ctypes.c_int64),
        ("ms_kernel", ctypes.c_double),
        ("wave_steps", ctypes.c_int64),
        ("lane_steps", ctypes.c_int64),
    ]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["active_per_round"] = list(self.active_per_round)
        return d


class TreeDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nclouds", ctypes.c_int), ("leaf_size", ctypes.c_int),
                ("bbox", ctypes.c_float * 6), ("key_bits", ctypes.c_int),
                ("axis_bits", ctypes.c_int)]


_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_flt = ctypes.c_float

# (name, restype, argtypes) of every entry point in include/gprt_hip.h
SIGNATURES = [
    ("gprt_hip_version", _int, []),
    ("gprt_hip_last_error", ctypes.c_char_p, []),
    ("gprt_hip_clear_error", None, []),
    ("gprt_hip_debug_set_query_sort_bits", None, [_int]),
    ("gprt_hip_debug_alloc_stats", None, [ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    ("gprt_hip_debug_fail_alloc", None, [_i64]),
    ("gprt_hip_debug_stream_copy", _int, [_vp, _vp, _vp, ctypes.c_size_t]),
    ("gprt_hip_knn_noalloc", _int, [_vp, _vp, _vp, _i64, _int, ctypes.POINTER(Config), _vp, _vp]),
    ("gprt_hip_knn_noalloc_max_queries", _i64, []),
    ("gprt_hip_config_default", None, [ctypes.POINTER(Config)]),
    ("gprt_hip_ctx_create", _int, [_int, ctypes.POINTER(_vp)]),
    ("gprt_hip_ctx_destroy", _int, [_vp]),
    ("gprt_hip_ctx_sync", _int, [_vp]),
    ("gprt_hip_stats", _int, [_vp, ctypes.POINTER(RunStats)]),
    ("gprt_hip_num_devices", _int, [ctypes.POINTER(_int)]),
    ("gprt_hip_device_alloc", _int, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    ("gprt_hip_device_free", _int, [_vp, _vp]),
    ("gprt_hip_copy", _int, [_vp, _vp, _vp, ctypes.c_size_t, _int]),
    ("gprt_hip_tree_build", _int, [_vp, _vp, _i64, _int, _int, ctypes.POINTER(_vp)]),
    ("gprt_hip_tree_build_batched", _int, [_vp, _vp, _vp, _int, _int, _int, ctypes.POINTER(_vp)]),
    ("gprt_hip_tree_destroy", _int, [_vp]),
    ("gprt_hip_tree_info", _int, [_vp, ctypes.POINTER(TreeDesc)]),
    ("gprt_hip_tree_export", _int, [_vp, _vp, _vp, _vp, _vp]),
    ("gprt_hip_knn", _int, [_vp, _vp, _vp, _i64, _int, ctypes.POINTER(Config), _vp, _vp, _int]),
    ("gprt_hip_knn_plain", _int, [_vp, _vp, _vp, _i64, _int, _vp, _vp, _int]),
    ("gprt_hip_knn_batched", _int, [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(Config), _vp, _vp, _int]),
    ("gprt_hip_range", _int, [_vp, _vp, _vp, _i64, _flt, _vp, _vp, _i64, _int]),
    ("gprt_hip_bh_forces", _int, [_vp, _vp, _i64, _flt, _flt, _flt, _int, _vp, _int]),
    ("gprt_hip_morton_codes", _int, [_vp, _vp, _i64, _vp, _vp, _int]),
    ("gprt_hip_sample_radius", _int, [_vp, _vp, _i64, _vp, _i64, _int, _int, _flt, _int,
                                      ctypes.POINTER(_flt), _int]),
    ("gprt_hip_sort_pairs", _int, [_vp, _vp, _vp, _i64, _int, _int, _int]),
    ("gprt_hip_autotune", _int, [_vp, _vp, _i64, _vp, _i64, _int, _int, ctypes.POINTER(Config),
                                 ctypes.POINTER(_int), ctypes.c_char_p, _int]),
    ("gprt_hip_autotune_ex", _int, [_vp, _vp, _i64, _vp, _i64, _int, _int, ctypes.POINTER(Config),
                                    ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.c_char_p, _int]),
    ("gprt_hip_autotune_batched", _int, [_vp, _vp, _vp, _int, _vp, _vp, _vp, _int, ctypes.POINTER(Config),
                                         ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.c_char_p, _int]),
    ("gprt_hip_read_points", _int, [ctypes.c_char_p, _int, ctypes.c_void_p, ctypes.c_int64,
                                    ctypes.POINTER(ctypes.c_int64)]),
    ("gprt_hip_schedule_from_json", _int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(Config),
                                           ctypes.POINTER(_int)]),
    ("gprt_hip_schedule_from_json_ex", _int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(Config),
                                              ctypes.POINTER(_int), ctypes.POINTER(_int)]),
    ("gprt_hip_codegen", _int, [ctypes.c_char_p, ctypes.c_size_t, _vp, ctypes.c_size_t,
                                ctypes.POINTER(ctypes.c_size_t)]),
    ("gprt_hip_codegen_compile", _int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p,
                                        ctypes.POINTER(ctypes.c_size_t)]),
    # the reference FFI (gprt-optix/src/lib.rs:13-31)
    ("gprt_pipeline_create", _vp, [_vp, ctypes.c_size_t, _int]),
    ("gprt_pipeline_destroy", None, [_vp]),
    ("gprt_bvh_build", _vp, [_vp, _vp, _vp, _int, _int]),
    ("gprt_bvh_build_triangles", _vp, [_vp, _vp, _vp, _int, _vp]),
    ("gprt_bvh_refit", None, [_vp, _vp, _flt, _int]),
    ("gprt_bvh_destroy", None, [_vp]),
    ("gprt_register_array", None, [_vp, ctypes.c_char_p, ctypes.c_uint, _int]),
    ("gprt_register_value", None, [_vp, ctypes.c_char_p]),
    ("gprt_execute", None, [_vp, _vp, _vp, _int]),
    ("gprt_retrieve_array_lengths", None, [_vp, ctypes.c_char_p, _vp, _int]),
    ("gprt_retrieve_array_flat", None, [_vp, ctypes.c_char_p, _vp, _int]),
    ("gprt_retrieve_value", None, [_vp, ctypes.c_char_p, _vp]),
    ("gprt_execute_autorope_soa_ffi", None, [_vp, _vp, _int, _vp, _int, _vp, _flt, _vp]),
]

_lib = None


def lib():
    """Load libgprt_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python rt-dsl_amd/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            if not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        raise GprtError(rc, lib().gprt_hip_last_error().decode(errors="replace"))


def default_config(**kw):
    c = Config()
    lib().gprt_hip_config_default(ctypes.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError(k)
        setattr(c, k, v)
    return c


class DeviceArray:
    """A C-contiguous array resident in the context device's HBM, allocated
    through the library (gprt_hip_device_alloc).  Torch is not used on the GPU
    side: its wheel carries its own HIP runtime, which cannot share a process
    with the system ROCm runtime this library links."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(x) for x in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(lib().gprt_hip_device_alloc(ctx.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, ctx, a):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        _check(lib().gprt_hip_copy(ctx.h, d.ptr, a.ctypes.data, d.nbytes, 0))
        return d

    def to_host(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib().gprt_hip_copy(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes, 1))
        return out

    def __len__(self):
        return self.shape[0]

    def free(self):
        if self.ptr and self.ctx.h:
            lib().gprt_hip_device_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _is_dev(a):
    return isinstance(a, DeviceArray)


def _ptr(a, dtype=None):
    """Pointer + device flag of a numpy array or a DeviceArray."""
    if a is None:
        return None, None
    if _is_dev(a):
        return a.ptr, True
    if dtype is not None and a.dtype != dtype:
        raise TypeError(f"expected {dtype}, got {a.dtype}")
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("host arrays must be C-contiguous")
    return a.ctypes.data, False


def _points(a, cols=3):
    if _is_dev(a):
        if a.dtype != np.float32 or len(a.shape) != 2 or a.shape[1] != cols:
            raise ValueError(f"expected an (n, {cols}) float32 device array")
        return a
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != cols:
        raise ValueError(f"expected an (n, {cols}) float32 array")
    return a


class Context:
    """One HIP device + stream (gprt_hip_ctx)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().gprt_hip_ctx_create(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().gprt_hip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(lib().gprt_hip_ctx_sync(self.h))

    def stats(self):
        s = RunStats()
        _check(lib().gprt_hip_stats(self.h, ctypes.byref(s)))
        return s.as_dict()

    # ---- trees ----
    def build(self, pts, leaf_size=8, ref_keys=False, sah=False):
        """LBVH over pts; ref_keys sorts by the reference's 30-bit Morton codes;
        sah splits the Morton-ordered points by binned SAH instead of the Karras
        hierarchy (GPRT_BUILD_SAH)."""
        pts = _points(pts)
        p, dev = _ptr(pts)
        t = ctypes.c_void_p()
        flags = (PTR_DEVICE if dev else PTR_HOST) | (KEYS_REF30 if ref_keys else 0) | (BUILD_SAH if sah else 0)
        _check(lib().gprt_hip_tree_build(self.h, p, len(pts), leaf_size, flags, ctypes.byref(t)))
        return Tree(self, t, pts)

    def build_batched(self, pts, cloud_offsets, leaf_size=8, ref_keys=False, sah=False):
        pts = _points(pts)
        off = np.ascontiguousarray(cloud_offsets, np.int64)
        p, dev = _ptr(pts)
        t = ctypes.c_void_p()
        flags = (PTR_DEVICE if dev else PTR_HOST) | (KEYS_REF30 if ref_keys else 0) | (BUILD_SAH if sah else 0)
        _check(lib().gprt_hip_tree_build_batched(self.h, p, off.ctypes.data, len(off) - 1, leaf_size, flags,
                                                 ctypes.byref(t)))
        return Tree(self, t, pts)

    # ---- pieces ----
    def morton_codes(self, pts):
        pts = _points(pts)
        codes = np.empty(len(pts), np.uint32)
        bbox = np.empty(6, np.float32)
        p, dev = _ptr(pts)
        if dev:
            raise ValueError("morton_codes takes host arrays")
        _check(lib().gprt_hip_morton_codes(self.h, p, len(pts), codes.ctypes.data, bbox.ctypes.data, PTR_HOST))
        return codes, bbox

    def sample_radius(self, pts, queries, sampler=SAMPLER_LCG, heuristic=RADIUS_PERCENTILE, param=0.95, k=5):
        pts, queries = _points(pts), _points(queries)
        r = ctypes.c_float()
        _check(lib().gprt_hip_sample_radius(self.h, pts.ctypes.data, len(pts), queries.ctypes.data, len(queries),
                                            sampler, heuristic, param, k, ctypes.byref(r), PTR_HOST))
        return float(np.float32(r.value))

    def sort_pairs(self, keys, vals, begin_bit=0, end_bit=64):
        keys = np.ascontiguousarray(keys, np.uint64).copy()
        vals = np.ascontiguousarray(vals, np.uint32).copy()
        _check(lib().gprt_hip_sort_pairs(self.h, keys.ctypes.data, vals.ctypes.data, len(keys), begin_bit,
                                         end_bit, PTR_HOST))
        return keys, vals

    # ---- Barnes-Hut ----
    def barnes_hut(self, pos_mass, theta=0.5, eps=1e-3, G=1.0, leaf_size=8, out=None, law="plummer", flags=0):
        """Accelerations (n, 3) of float4 bodies (x, y, z, m).  law "plummer":
        G m_j d / (|d|^2 + eps^2)^1.5; law "reference": the reference shader's
        G m_j d / (|d|^2 sqrt(|d|^2 + eps^2)) over pairs with |d|^2 > 1e-10."""
        if law not in ("plummer", "reference"):
            raise ValueError("law is 'plummer' or 'reference'")
        pm = _points(pos_mass, 4)
        p, dev = _ptr(pm)
        n = len(pm)
        if out is None:
            out = DeviceArray(self, (n, 3), np.float32) if dev else np.empty((n, 3), np.float32)
        o, _ = _ptr(out)
        flags |= (PTR_DEVICE if dev else PTR_HOST) | (BH_REFERENCE_LAW if law == "reference" else 0)
        _check(lib().gprt_hip_bh_forces(self.h, p, n, theta, eps, G, leaf_size, o, flags))
        return out

    def autotune(self, pts, queries, k, budget_trials=24, json_path=None, with_tree=False):
        """The project's autotuner (gprt_autotune! of the reference): returns
        (Config, leaf_size) and optionally writes the Schedule-superset JSON;
        with_tree also chooses between the LBVH and the SAH tree and returns
        (Config, leaf_size, sah)."""
        pts, queries = _points(pts), _points(queries)
        p, dev = _ptr(pts)
        qp, qdev = _ptr(queries)
        if dev != qdev:
            raise ValueError("points and queries must both be host arrays or both device arrays")
        cfg = Config()
        leaf = ctypes.c_int()
        bflags = ctypes.c_int()
        _check(lib().gprt_hip_autotune_ex(self.h, p, len(pts), qp, len(queries), k, budget_trials,
                                          ctypes.byref(cfg), ctypes.byref(leaf),
                                          ctypes.byref(bflags) if with_tree else None,
                                          json_path.encode() if json_path else None,
                                          PTR_DEVICE if dev else PTR_HOST))
        if with_tree:
            return cfg, leaf.value, bool(bflags.value & BUILD_SAH)
        return cfg, leaf.value

    def autotune_batched(self, pts, cloud_offsets, queries, query_offsets, k_per_cloud, budget_trials=24,
                         json_path=None, with_tree=False):
        """The autotuner over an uneven batch (gprt_hip_autotune_batched):
        (Config, leaf_size[, sah]) for gprt_hip_knn_batched on these clouds."""
        pts, queries = _points(pts), _points(queries)
        p, dev = _ptr(pts)
        qp, qdev = _ptr(queries)
        if dev != qdev:
            raise ValueError("points and queries must both be host arrays or both device arrays")
        poff = np.ascontiguousarray(cloud_offsets, np.int64)
        qoff = np.ascontiguousarray(query_offsets, np.int64)
        kc = np.ascontiguousarray(k_per_cloud, np.int32)
        if not (len(poff) == len(qoff) == len(kc) + 1):
            raise ValueError("need nclouds + 1 cloud and query offsets and nclouds k values")
        cfg = Config()
        leaf = ctypes.c_int()
        bflags = ctypes.c_int()
        _check(lib().gprt_hip_autotune_batched(self.h, p, poff.ctypes.data, len(kc), qp, qoff.ctypes.data,
                                               kc.ctypes.data, budget_trials, ctypes.byref(cfg),
                                               ctypes.byref(leaf), ctypes.byref(bflags) if with_tree else None,
                                               json_path.encode() if json_path else None,
                                               PTR_DEVICE if dev else PTR_HOST))
        if with_tree:
            return cfg, leaf.value, bool(bflags.value & BUILD_SAH)
        return cfg, leaf.value


class Tree:
    """A GPU LBVH (gprt_hip_tree)."""

    def __init__(self, ctx, h, pts_ref):
        self.ctx = ctx
        self.h = h
        self._keep = pts_ref  # device tensors must outlive nothing here, but keep for symmetry
        d = TreeDesc()
        _check(lib().gprt_hip_tree_info(h, ctypes.byref(d)))
        self.n = d.n
        self.nclouds = d.nclouds
        self.leaf_size = d.leaf_size
        self.bbox = np.array(list(d.bbox), np.float32)
        self.key_bits = d.key_bits
        self.axis_bits = d.axis_bits

    def info(self):
        return {"n": self.n, "nclouds": self.nclouds, "leaf_size": self.leaf_size, "bbox": self.bbox,
                "key_bits": self.key_bits, "axis_bits": self.axis_bits}

    def close(self):
        if self.h:
            lib().gprt_hip_tree_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def export(self):
        n = self.n
        keys = np.empty(n, np.uint64)
        perm = np.empty(n, np.uint32)
        nodes = np.empty((max(n - 1, 0), 16), np.float32)
        roots = np.empty((self.nclouds, 2), np.int32)
        _check(lib().gprt_hip_tree_export(self.h, keys.ctypes.data, perm.ctypes.data,
                                          nodes.ctypes.data if n > 1 else None, roots.ctypes.data))
        return keys, perm, nodes, roots

    def _out(self, queries, width, dtype_np):
        nq = len(queries)
        if _is_dev(queries):
            return DeviceArray(self.ctx, (nq, width), dtype_np)
        return np.empty((nq, width), dtype_np)

    def k_nn(self, queries, k, config=None, with_d2=False, ids_out=None, d2_out=None, sync=True):
        """Batched FP32 + FP64-recheck exact k-NN; ids (nq, k) ordered by (d64, id)."""
        queries = _points(queries)
        cfg = config if config is not None else default_config()
        if ids_out is None:
            ids_out = self._out(queries, k, np.uint32)
        if with_d2 and d2_out is None:
            d2_out = self._out(queries, k, np.float64)
        q, dev = _ptr(queries)
        i, _ = _ptr(ids_out)
        d, _ = _ptr(d2_out) if with_d2 else (None, None)
        flags = (PTR_DEVICE if dev else PTR_HOST) | (0 if sync or not dev else NO_SYNC)
        _check(lib().gprt_hip_knn(self.ctx.h, self.h, q, len(queries), k, ctypes.byref(cfg), i, d, flags))
        return (ids_out, d2_out) if with_d2 else ids_out

    def k_nn_plain(self, queries, k, with_d2=False, ids_out=None, d2_out=None, sync=True, counters=False):
        """Plain single-launch FP64 form (input order, r = inf)."""
        queries = _points(queries)
        if ids_out is None:
            ids_out = self._out(queries, k, np.uint32)
        if with_d2 and d2_out is None:
            d2_out = self._out(queries, k, np.float64)
        q, dev = _ptr(queries)
        i, _ = _ptr(ids_out)
        d, _ = _ptr(d2_out) if with_d2 else (None, None)
        flags = (PTR_DEVICE if dev else PTR_HOST) | (0 if sync or not dev else NO_SYNC)
        if counters:
            flags |= COLLECT_COUNTERS
        _check(lib().gprt_hip_knn_plain(self.ctx.h, self.h, q, len(queries), k, i, d, flags))
        return (ids_out, d2_out) if with_d2 else ids_out

    def k_nn_batched(self, queries, query_offsets, k_per_cloud, config=None, with_d2=False):
        """Per-cloud query sets with per-cloud k; returns the flat id array and the
        output offsets (out_offsets[c] = sum of nq_c' * k_c' for c' < c)."""
        queries = _points(queries)
        qoff = np.ascontiguousarray(query_offsets, np.int64)
        kc = np.ascontiguousarray(k_per_cloud, np.int32)
        counts = np.diff(qoff)
        out_off = np.concatenate([[0], np.cumsum(counts * kc)]).astype(np.int64)
        ids = np.empty(int(out_off[-1]), np.uint32)
        d2 = np.empty(int(out_off[-1]), np.float64) if with_d2 else None
        cfg = config if config is not None else default_config()
        q, dev = _ptr(queries)
        if dev:
            raise ValueError("k_nn_batched takes host arrays")
        _check(lib().gprt_hip_knn_batched(self.ctx.h, self.h, q, qoff.ctypes.data, kc.ctypes.data,
                                          ctypes.byref(cfg), ids.ctypes.data,
                                          d2.ctypes.data if with_d2 else None, PTR_HOST))
        return (ids, out_off, d2) if with_d2 else (ids, out_off)

    def range_query(self, queries, radius, flags=0, stats=None):
        """All ids with d64 <= r*r per query, CSR (offsets, ids), ordered by (d64, id).
        flags (COLLECT_COUNTERS, TIME_KERNELS) apply to both calls; with a list
        `stats`, the statistics of the counting call and of the filling call
        are appended to it."""
        queries = _points(queries)
        q, dev = _ptr(queries)
        if dev:
            raise ValueError("range_query takes host arrays")
        nq = len(queries)
        off = np.zeros(nq + 1, np.int64)
        _check(lib().gprt_hip_range(self.ctx.h, self.h, q, nq, radius, off.ctypes.data, None, 0, PTR_HOST | flags))
        if stats is not None:
            stats.append(self.ctx.stats())
        ids = np.empty(max(int(off[-1]), 1), np.uint32)
        _check(lib().gprt_hip_range(self.ctx.h, self.h, q, nq, radius, off.ctypes.data, ids.ctypes.data,
                                    len(ids), PTR_HOST | flags))
        if stats is not None:
            stats.append(self.ctx.stats())
        return off, ids[: int(off[-1])]


CSV_POINTS = 0
CSV_BODIES = 1


def read_points(path, fmt=CSV_POINTS):
    """(n, 3) float32 points (CSV_POINTS) or (n, 4) x, y, z, mass bodies
    (CSV_BODIES) from a text file, read as the reference's sandbox reads it."""
    width = 4 if fmt == CSV_BODIES else 3
    # a row takes at least 6 (points: "a,b,c\n") or 5 (bodies: ",,,m\n") bytes
    cap = (os.path.getsize(path) + 1) // (6 if fmt == CSV_POINTS else 5) + 1
    out = np.empty((cap, width), np.float32)
    rows = ctypes.c_int64()
    _check(lib().gprt_hip_read_points(os.fsencode(path), fmt, out.ctypes.data, cap, ctypes.byref(rows)))
    return out[:rows.value].copy()


def parse_schedule(text, with_tree=False):
    """(Config, leaf_size) from the text of a Schedule / wisdom JSON object;
    with_tree adds sah (build_strategy PreferFastTrace, the default)."""
    raw = text.encode() if isinstance(text, str) else bytes(text)
    cfg = Config()
    leaf = _int()
    bflags = _int()
    _check(lib().gprt_hip_schedule_from_json_ex(raw, len(raw), ctypes.byref(cfg), ctypes.byref(leaf),
                                                ctypes.byref(bflags)))
    if with_tree:
        return cfg, leaf.value, bool(bflags.value & BUILD_SAH)
    return cfg, leaf.value


def load_schedule(path, with_tree=False):
    """(Config, leaf_size[, sah]) from a Schedule / wisdom JSON file."""
    with open(path, "rb") as f:
        return parse_schedule(f.read(), with_tree)


def codegen(program):
    """HIP source of a shader-IR program (a dict or JSON text); host only."""
    import json as _json
    text = program if isinstance(program, str) else _json.dumps(program)
    raw = text.encode()
    n = ctypes.c_size_t()
    _check(lib().gprt_hip_codegen(raw, len(raw), None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value)
    _check(lib().gprt_hip_codegen(raw, len(raw), buf, n.value, ctypes.byref(n)))
    return buf.value.decode()


def codegen_compile(program, arch="gfx950"):
    """Lower and compile a shader-IR program with hiprtc; returns the code-object size."""
    import json as _json
    text = program if isinstance(program, str) else _json.dumps(program)
    raw = text.encode()
    n = ctypes.c_size_t()
    _check(lib().gprt_hip_codegen_compile(raw, len(raw), arch.encode(), ctypes.byref(n)))
    return n.value


def num_devices():
    n = ctypes.c_int()
    _check(lib().gprt_hip_num_devices(ctypes.byref(n)))
    return n.value


# ---- reference-style one-shot helpers (gprt-macros/src/lib.rs) ----
def k_nn(ctx, data, queries, k, config=None, leaf_size=8):
    """k_nn!(data, queries, k, out[, schedule]): exact ids, (nq, k)."""
    t = ctx.build(data, leaf_size)
    try:
        return t.k_nn(queries, k, config)
    finally:
        t.close()


def range_query(ctx, data, queries, radius, leaf_size=8):
    """r_nn! / range_query!: per-query neighbours within radius, CSR."""
    t = ctx.build(data, leaf_size)
    try:
        return t.range_query(queries, radius)
    finally:
        t.close()


def barnes_hut(ctx, pos_mass, theta=0.5, G=1.0, eps=1e-3, law="reference"):
    """barnes_hut!(bodies, theta, G, out): accelerations (n, 3), by default
    with the reference's force law (gprt-macros/src/lib.rs:785)."""
    return ctx.barnes_hut(pos_mass, theta, eps, G, law=law)
