"""Shader-IR programs: lowering to HIP and hiprtc compilation (host only).

The IR is the reference's ShaderNode tree (gprt-ir/src/lib.rs:8-18) in JSON;
the lowering follows gprt-codegen's lower_node (gprt-codegen/src/lib.rs:78-119)
with HIP in place of CUDA/OptiX.  Launches are in test_gpu_codegen.py."""
import json

import pytest

import gprt_hip

PUSH_EVEN = {
    "raygen": {"TraceRay": {}},
    "intersection": {"If": {"condition": "gprt_dist_sq <= __uint_as_float(p2) && (gprt_prim % 2u) == 0u",
                            "then": "ReportIntersection"}},
    "anyhit": {"Block": [{"PushToDynamicArray": {"array": 0, "value": "gprt_prim"}}, "IgnoreIntersection"]},
}


def test_lowering_of_every_node():
    prog = {
        "raygen": {"Block": [
            {"UpdatePayload": {"register": 3, "value": "7u"}},
            {"TraceRay": {"on_hit": {"Raw": "p0 = 1u;"}, "on_miss": {"Raw": "p0 = 2u;"}}},
            {"If": {"condition": "p0 == 1u", "then": {"PushToDynamicArray": {"array": 1, "value": "p3"}},
                    "else": {"Raw": "p1 = 0u;\np2 = 0u;"}}},
        ]},
        "intersection": {"If": {"condition": "gprt_dist_sq < 1.0f", "then": "ReportIntersection"}},
        "anyhit": {"Block": ["IgnoreIntersection", "TerminateRay"]},
    }
    src = gprt_hip.codegen(prog)
    assert 'extern "C" __global__' in src and "gprt_generated(GprtLaunch params)" in src
    assert "p3 = (unsigned int)(7u);" in src
    assert "gprt_walk(params, gprt_ray, p2, gprt_stop, gprt_leaf);" in src
    assert "if (gprt_dist_sq < 1.0f)" in src and "gprt_reported = true;" in src
    assert "gprt_ignore = true;" in src and "gprt_terminate = true;" in src
    assert "if (p0 == 1u)" in src and "else" in src
    assert "bundle->dyn_lens[1]" in src and "bundle->dyn_caps[1]" in src and "(unsigned int)(p3)" in src
    assert "p1 = 0u;" in src and "p2 = 0u;" in src  # a Raw block keeps its lines
    # on_hit / on_miss follow the walk
    assert src.index("gprt_walk(") < src.index("p0 = 1u;") < src.index("p0 = 2u;")


@pytest.mark.parametrize("prog,msg", [
    ("{", "expected a string"),
    ('{"raygen": {"TraceRay": {}}} x', "trailing"),
    ('{"anyhit": "IgnoreIntersection"}', "raygen program is required"),
    ('{"raygen": "Teleport"}', "unknown node"),
    ('{"raygen": "ReportIntersection"}', "outside intersection"),
    ('{"raygen": {"TraceRay": {}}, "intersection": {"TraceRay": {}}}', "TraceRay outside raygen"),
    ('{"raygen": {"PushToDynamicArray": {"array": 8, "value": "1"}}}', "array must be an integer"),
    ('{"raygen": {"UpdatePayload": {"register": 4, "value": "1"}}}', "register must be an integer"),
    ('{"raygen": {"If": {"then": "TerminateRay"}}}', "condition must be"),
    ('{"raygen": {"TraceRay": {}}, "shadow": "TerminateRay"}', "unknown program"),
])
def test_lowering_rejects(prog, msg):
    with pytest.raises(gprt_hip.GprtError, match=msg):
        gprt_hip.codegen(prog)


def test_hiprtc_compiles_for_gfx950():
    assert gprt_hip.codegen_compile(PUSH_EVEN) > 1000


def test_hiprtc_reports_compile_errors():
    bad = {"raygen": {"Raw": "this is not HIP;"}}
    with pytest.raises(gprt_hip.GprtError, match="hiprtc"):
        gprt_hip.codegen_compile(json.dumps(bad))
