"""Shader-IR programs compiled at pipeline creation and run through the
reference FFI (gprt_pipeline_create / gprt_register_array / gprt_execute /
gprt_retrieve_*), checked against numpy restatements of what the programs
compute.  The FP32 distance of the generated walk is (dx*dx + dy*dy) + dz*dz
with every operation rounded, which numpy float32 arithmetic reproduces."""
import json

import numpy as np
import pytest

import gprt_hip
from gprt_hip import datasets as D

pytestmark = pytest.mark.gpu


def _d32(pts, q):
    d = pts[None, :, :] - q[:, None, :]  # float32, rounded per operation
    return (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]


def _pipe(program):
    L = gprt_hip.lib()
    ir = b"gprt-hip:ir:" + json.dumps(program).encode()
    p = L.gprt_pipeline_create(ir, len(ir), 0)
    assert p, L.gprt_hip_last_error()
    return L, p


def _run(L, p, pts, q, r, arrays):
    geom = np.concatenate([pts, np.full((len(pts), 1), r, np.float32)], 1)
    bvh = L.gprt_bvh_build(p, geom.ctypes.data, None, len(pts), geom.nbytes)
    assert bvh
    for name, cap in arrays:
        L.gprt_register_array(p, name, cap, len(q))
    qq = np.concatenate([q, np.full((len(q), 1), r, np.float32)], 1)
    L.gprt_execute(p, bvh, qq.ctypes.data, len(q))
    out = {}
    for name, cap in arrays:
        lens = np.empty(len(q), np.uint32)
        flat = np.empty(len(q) * cap, np.uint32)
        L.gprt_retrieve_array_lengths(p, name, lens.ctypes.data, len(q))
        L.gprt_retrieve_array_flat(p, name, flat.ctypes.data, flat.size)
        out[name] = (lens, flat.reshape(len(q), cap))
    L.gprt_bvh_destroy(bvh)
    return out


def test_generated_predicate_push():
    """A spatial predicate beyond d <= r: points within r whose id is even."""
    prog = {
        "raygen": {"TraceRay": {}},
        "intersection": {"If": {"condition": "gprt_dist_sq <= __uint_as_float(p2) && (gprt_prim % 2u) == 0u",
                                "then": "ReportIntersection"}},
        "anyhit": {"Block": [{"PushToDynamicArray": {"array": 0, "value": "gprt_prim"}}, "IgnoreIntersection"]},
    }
    L, p = _pipe(prog)
    try:
        pts, q = D.uniform_cloud(20_000, 1_000, seed=21)
        r = np.float32(0.04)
        cap = 64
        lens, flat = _run(L, p, pts, q, r, [(b"hits", cap)])[b"hits"]
        d = _d32(pts, q)
        want = (d <= np.float32(r * r)) & ((np.arange(len(pts)) % 2) == 0)[None, :]
        assert np.array_equal(lens, want.sum(1).astype(np.uint32))
        assert lens.max() <= cap
        for i in range(len(q)):
            assert sorted(flat[i, :lens[i]].tolist()) == np.nonzero(want[i])[0].tolist()
    finally:
        L.gprt_pipeline_destroy(p)


HEAP = {
    # the register-heap k-NN of gprt_rnn_run! (gprt-macros/src/lib.rs:305-377)
    # written against this device API: ids[0..K) then FP32 dists[K..2K), the
    # worst distance in p2 bounding the walk
    "raygen": {"Block": [
        {"Raw": "unsigned int* heap = (unsigned int*)bundle->dyn_ptrs[0] + (size_t)gprt_qid * bundle->dyn_caps[0];\n"
                "const unsigned int K = bundle->dyn_caps[0] / 2u;\n"
                "for (unsigned int i = 0; i < K; ++i) { heap[i] = 0xFFFFFFFFu; heap[K + i] = p2; }"},
        {"TraceRay": {}},
        {"Raw": "unsigned int cnt = 0;\n"
                "for (unsigned int i = 0; i < K; ++i) cnt += heap[i] != 0xFFFFFFFFu;\n"
                "((unsigned int*)bundle->dyn_lens[0])[gprt_qid] = cnt;"},
    ]},
    "intersection": {"If": {"condition": "gprt_dist_sq <= __uint_as_float(p2)", "then": "ReportIntersection"}},
    "anyhit": {"Block": [
        {"Raw": "if (gprt_dist_sq < __uint_as_float(heap[2u * K - 1u])) {\n"
                "    heap[K - 1u] = gprt_prim;\n"
                "    heap[2u * K - 1u] = __float_as_uint(gprt_dist_sq);\n"
                "    for (int i = (int)K - 2; i >= 0; --i) {\n"
                "        const float dc = __uint_as_float(heap[K + i]), dn = __uint_as_float(heap[K + i + 1]);\n"
                "        if (dc > dn) {\n"
                "            heap[K + i] = __float_as_uint(dn); heap[K + i + 1] = __float_as_uint(dc);\n"
                "            const unsigned int t = heap[i]; heap[i] = heap[i + 1]; heap[i + 1] = t;\n"
                "        } else break;\n"
                "    }\n"
                "    p2 = heap[2u * K - 1u];\n"
                "}"},
        "IgnoreIntersection",
    ]},
}


def test_generated_register_heap_knn():
    L, p = _pipe(HEAP)
    try:
        pts, q = D.uniform_cloud(20_000, 1_000, seed=22)
        r = np.float32(0.05)
        k = 6
        lens, flat = _run(L, p, pts, q, r, [(b"knn", 2 * k)])[b"knn"]
        d = _d32(pts, q)
        r2 = np.float32(r * r)
        for i in range(len(q)):
            cand = np.nonzero(d[i] < r2)[0]
            order = cand[np.lexsort((cand, d[i][cand]))][:k]
            m = len(order)
            assert lens[i] == m
            # without FP32 ties among the first k + 1 the heap's order is the distance order
            if len(np.unique(d[i][cand])) == len(cand):
                assert flat[i, :m].tolist() == order.tolist(), i
            else:  # ties: the heap keeps the first visited of equal distances
                assert np.array_equal(np.sort(d[i][flat[i, :m]]), np.sort(d[i][order])), i
            assert np.all(flat[i, m:k] == 0xFFFFFFFF)
            assert np.array_equal(flat[i, k:k + m].view(np.float32), d[i][order])
    finally:
        L.gprt_pipeline_destroy(p)


def test_generated_program_compile_error():
    L = gprt_hip.lib()
    ir = b'gprt-hip:ir:{"raygen": {"Raw": "not_a_function();"}}'
    p = L.gprt_pipeline_create(ir, len(ir), 0)
    assert not p
    assert b"hiprtc" in L.gprt_hip_last_error()
