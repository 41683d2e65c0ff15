"""The reference's legacy generic trace API, over the library's drop-in FFI.

Mirrors gprt_build_index! (gprt-macros/src/lib.rs:221-237), gprt_trace!
(:625-672) and gprt_trace_batched! (:674-727), with IndexBuilder /
SpatialIndex (gprt-optix/src/lib.rs:164-187), by driving the same 13 symbols a
Rust crate binds (gprt_pipeline_create with the "gprt-hip:push" program,
gprt_bvh_build, gprt_register_array, gprt_execute, gprt_retrieve_*).

A ray is (x, y, z, r): its origin and search radius (Ray::query,
gprt-core/src/lib.rs:52).  It hits the sphere primitives whose centre c has
d64(origin, c) <= r * r, the binary64 contract of the library; the spheres'
own radii bound nothing beyond that (in the reference they size the AABBs the
hardware walk tests, and the any-hit test compares against the ray's r).
Within one ray the hits come in (d64, id) order, the order the reference leaves
to its any-hit invocations.

- trace(index, rays): every hit of every ray in one list, ray by ray, at most
  1,000,000 ids (gprt_trace!'s one shared array, :662-670);
- trace_batched(index, rays, cap): per-ray slots (flat, nq * cap), the hit
  count of each ray (which may exceed cap; only cap are kept) and cap, the
  tuple gprt_trace_batched! returns (:722-724); unused slots hold 0.

Triangle primitives (gprt-core/src/lib.rs:96-121) are not traced: the
reference builds them as plain AABBs with an empty intersection program
(gprt-optix/host/main.cpp, gprt_bvh_build_triangles), so its GPU path reports
no triangle hits, and nothing in it calls the CPU Triangle::intersect.
"""
import ctypes

import numpy as np

import gprt_hip

FLAT_CAP = 1_000_000  # gprt_trace!'s array capacity (gprt-macros/src/lib.rs:662)
_PUSH = b"gprt-hip:push"


def _call(fn, *args):
    """Run one no-return-value FFI symbol and raise on the error it leaves."""
    L = gprt_hip.lib()
    L.gprt_hip_clear_error()
    r = fn(*args)
    msg = L.gprt_hip_last_error()
    if msg:
        raise gprt_hip.GprtError(gprt_hip.ERR_STATE, msg.decode(errors="replace"))
    return r


class IndexBuilder:
    """A scene of spheres (x, y, z, radius) and its schedule; the BVH is built
    on the first trace and rebuilt after mark_dirty()."""

    def __init__(self, spheres, schedule=None):
        self.spheres = np.ascontiguousarray(spheres, np.float32).reshape(-1, 4)
        self.schedule = schedule
        L = gprt_hip.lib()
        L.gprt_hip_clear_error()
        self._pipe = L.gprt_pipeline_create(_PUSH, len(_PUSH), 0)
        if not self._pipe:
            raise gprt_hip.GprtError(gprt_hip.ERR_STATE, L.gprt_hip_last_error().decode(errors="replace"))
        self._bvh = None

    def mark_dirty(self):
        self._drop_bvh()

    def _drop_bvh(self):
        if self._bvh:
            gprt_hip.lib().gprt_bvh_destroy(self._bvh)
            self._bvh = None

    def _ensure_bvh(self):
        if self._bvh is None:
            s = self.spheres
            aabbs = np.concatenate([s[:, :3] - s[:, 3:], s[:, :3] + s[:, 3:]], 1).astype(np.float32)
            L = gprt_hip.lib()
            bvh = _call(L.gprt_bvh_build, self._pipe, s.ctypes.data, aabbs.ctypes.data, len(s), s.nbytes)
            if not bvh:
                raise gprt_hip.GprtError(gprt_hip.ERR_STATE, "gprt_bvh_build returned NULL")
            self._bvh = bvh
        return self._bvh

    def _run(self, name, rays, cap, nq_slots):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 4)
        L = gprt_hip.lib()
        bvh = self._ensure_bvh()
        _call(L.gprt_register_array, self._pipe, name, cap, nq_slots)
        _call(L.gprt_execute, self._pipe, bvh, rays.ctypes.data, len(rays))
        lens = np.zeros(nq_slots, np.uint32)
        flat = np.zeros(nq_slots * cap, np.uint32)
        _call(L.gprt_retrieve_array_lengths, self._pipe, name, lens.ctypes.data, nq_slots)
        _call(L.gprt_retrieve_array_flat, self._pipe, name, flat.ctypes.data, flat.size)
        return flat, lens

    def close(self):
        self._drop_bvh()
        if self._pipe:
            gprt_hip.lib().gprt_pipeline_destroy(self._pipe)
            self._pipe = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_index(data, geom=0.0, schedule=None):
    """gprt_build_index!(data, geom[, schedule]).  `geom` is the sphere radius
    for every point, or a function taking the (n, 3) points and returning
    (n, 4) spheres (the per-point closure of the macro, applied to all)."""
    pts = np.ascontiguousarray(data, np.float32).reshape(-1, 3)
    if callable(geom):
        spheres = np.asarray(geom(pts), np.float32).reshape(len(pts), 4)
    else:
        spheres = np.concatenate([pts, np.full((len(pts), 1), geom, np.float32)], 1)
    return IndexBuilder(spheres, schedule)


def trace(index, rays):
    """gprt_trace!(index, rays, |hit| out.push(hit)): the ids hit by all rays
    in one list, ray by ray, truncated at 1,000,000."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 4)
    if len(rays) == 0:
        return np.zeros(0, np.uint32)
    flat, lens = index._run(b"flat_out", rays, FLAT_CAP, 1)
    return flat[:min(int(lens[0]), FLAT_CAP)].copy()


def trace_batched(index, rays, cap):
    """gprt_trace_batched!(index, rays, cap) -> (slots, counts, cap)."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 4)
    cap = int(cap)
    if cap < 0 or cap > 0xFFFFFFFF:
        raise ValueError("cap must fit in 32 bits")
    if len(rays) == 0:
        return np.zeros(0, np.uint32), np.zeros(0, np.uint32), cap
    flat, lens = index._run(b"batched_out", rays, cap, len(rays))
    return flat, lens, cap
