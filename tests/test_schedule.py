"""Schedule / wisdom JSON (gprt-ir/src/lib.rs:58-89) read into a config (host only).

The fixtures tests/golden/wisdom_*.json are the reference's own wisdom files
(kitti_wisdom.json, 3droad_wisdom.json, 3diono_wisdom.json, porto_wisdom.json),
kept byte for byte as data.  They lack build_strategy and geom_type, which the
reference's serde reader requires; here missing fields take Schedule::default()
(gprt-ir/src/lib.rs:76-87)."""
import json
import os

import pytest

import gprt_hip

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# (file, SampledPercentile, radius_increment_mult), read from the fixtures' text
WISDOM = [
    ("wisdom_kitti.json", 0.5, 3.0),
    ("wisdom_3droad.json", 0.2, 2.0),
    ("wisdom_3diono.json", 0.02, 2.0),
    ("wisdom_porto.json", 0.1, 3.0),
]


@pytest.mark.parametrize("name,p,mult", WISDOM)
def test_reference_wisdom_files(name, p, mult):
    path = os.path.join(GOLDEN, name)
    raw = json.load(open(path))
    assert raw["radius_heuristic"] == {"SampledPercentile": p} and raw["radius_increment_mult"] == mult
    cfg, leaf = gprt_hip.load_schedule(path)
    assert cfg.radius_heuristic == gprt_hip.RADIUS_PERCENTILE
    assert cfg.radius_param == pytest.approx(p, rel=1e-7)
    assert cfg.radius_mult == mult
    assert cfg.sort_queries == 1  # use_morton_lbv: true
    assert leaf == 8
    # fields the file does not name keep the library's defaults
    d = gprt_hip.default_config()
    for f in ("sampler", "max_rounds", "traversal", "stack_depth", "guard_slots", "use_graph", "lds_stack",
              "compact_rounds", "grid_blocks"):
        assert getattr(cfg, f) == getattr(d, f), f


def test_schedule_defaults():
    """An empty object is Schedule::default(): SampledPercentile(0.10), mult 3.0,
    use_morton_lbv false, build_strategy PreferFastTrace (the SAH build)."""
    cfg, leaf = gprt_hip.parse_schedule("{}")
    assert cfg.radius_heuristic == gprt_hip.RADIUS_PERCENTILE
    assert cfg.radius_param == pytest.approx(0.10)
    assert cfg.radius_mult == 3.0 and cfg.sort_queries == 0 and leaf == 8
    assert gprt_hip.parse_schedule("{}", with_tree=True)[2] is True


def test_schedule_build_strategy():
    """build_strategy picks the tree: PreferFastTrace the SAH build,
    PreferFastBuild the LBVH; anything else is rejected."""
    assert gprt_hip.parse_schedule('{"build_strategy": "PreferFastTrace"}', with_tree=True)[2] is True
    assert gprt_hip.parse_schedule('{"build_strategy": "PreferFastBuild"}', with_tree=True)[2] is False
    for bad in ('{"build_strategy": "Fast"}', '{"build_strategy": 1}'):
        with pytest.raises(gprt_hip.GprtError):
            gprt_hip.parse_schedule(bad)


def test_schedule_heuristics():
    cfg, _ = gprt_hip.parse_schedule('{"radius_heuristic": "SampledMax", "use_morton_lbv": true}')
    assert cfg.radius_heuristic == gprt_hip.RADIUS_MAX and cfg.sort_queries == 1
    cfg, _ = gprt_hip.parse_schedule('{"radius_heuristic": {"Fixed": 0.25}, "radius_increment_mult": 1.5}')
    assert cfg.radius_heuristic == gprt_hip.RADIUS_FIXED and cfg.radius_param == 0.25 and cfg.radius_mult == 1.5


def test_schedule_full_reference_form():
    """All seven Schedule fields, as the reference writes them after tuning."""
    text = json.dumps({"radius_heuristic": {"SampledPercentile": 0.8}, "radius_increment_mult": 5.0,
                       "max_hits_per_query": 2000, "use_morton_lbv": True,
                       "memory_strategy": "PayloadRegisterHeap", "build_strategy": "PreferFastBuild",
                       "geom_type": "Triangles"})
    cfg, leaf, sah = gprt_hip.parse_schedule(text, with_tree=True)
    assert (cfg.radius_param, cfg.radius_mult, cfg.sort_queries, leaf) == (pytest.approx(0.8), 5.0, 1, 8)
    assert sah is False


def test_schedule_extension_object_round_trip():
    """The "gprt_hip" object (the form gprt_hip_autotune writes) restores
    every config field and the leaf size."""
    want = dict(sampler=gprt_hip.SAMPLER_STRIDED, radius_heuristic=gprt_hip.RADIUS_INFINITE, radius_param=0.75,
                radius_mult=2.5, max_rounds=9, traversal=gprt_hip.TRAV_WIDE, stack_depth=48, guard_slots=2,
                sort_queries=0, use_graph=0, grid_blocks=123, lds_stack=16, compact_rounds=1)
    text = json.dumps({"radius_increment_mult": 3.0, "use_morton_lbv": True,
                       "radius_heuristic": {"SampledPercentile": 0.1},
                       "gprt_hip": dict(want, leaf_size=16, ms_query=0.5)})
    cfg, leaf = gprt_hip.parse_schedule(text)
    assert leaf == 16
    for f, v in want.items():
        assert getattr(cfg, f) == pytest.approx(v), f


@pytest.mark.parametrize("bad", [
    "", "[1, 2]", "{", '{"radius_heuristic": "Median"}', '{"radius_heuristic": {"Percentile": 0.5}}',
    '{"use_morton_lbv": 1}', '{"gprt_hip": 3}', '{"gprt_hip": {"max_rounds": 1e12}}',
    '{"gprt_hip": {"stack_depth": 2.5}}', '{"radius_increment_mult": 1e300}',
    '{"radius_heuristic": {"Fixed": -1e39}}',
])
def test_schedule_rejects(bad):
    with pytest.raises(gprt_hip.GprtError) as e:
        gprt_hip.parse_schedule(bad)
    assert e.value.code == gprt_hip.ERR_ARG


def test_schedule_reads_only_the_given_bytes():
    """The reader stops at the given size: the input need not be NUL-terminated,
    and bytes after it (here digits) are not part of the number."""
    import ctypes
    text = b'{"radius_increment_mult": 2}' + b'5}'
    cfg = gprt_hip.Config()
    leaf = ctypes.c_int()
    L = gprt_hip.lib()
    assert L.gprt_hip_schedule_from_json(text, len(text) - 2, ctypes.byref(cfg), ctypes.byref(leaf)) == 0
    assert cfg.radius_mult == 2.0
    # cut inside the number: the object is unterminated
    assert L.gprt_hip_schedule_from_json(text, len(text) - 3, ctypes.byref(cfg), ctypes.byref(leaf)) == \
        gprt_hip.ERR_ARG


def test_schedule_rejects_bad_escape():
    with pytest.raises(gprt_hip.GprtError):
        gprt_hip.parse_schedule('{"\\uZZZZ": 1}')
    cfg, _ = gprt_hip.parse_schedule('{"\\u0072adius_increment_mult": 1.5}')  # r is "r"
    assert cfg.radius_mult == 1.5
