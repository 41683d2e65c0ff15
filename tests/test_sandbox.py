"""Sandbox harness pieces on the host: the text point reader (C ABI,
gprt_hip_read_points), argument parsing, schedule choice and output formats of
gprt-sandbox/src/main.rs and src/bin/barneshut.rs.  End-to-end runs are in
test_gpu_sandbox.py."""
import numpy as np
import pytest

import gprt_hip
from gprt_hip import barneshut, sandbox


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode() if isinstance(text, str) else text)
    return str(p)


# ---------------------------------------------------------------- reader
def test_read_points_rules(tmp_path):
    """io.rs:16-28: empty lines, lines with a bad or missing first-three field
    are dropped; extra fields, spaces and CRLF are fine; row order is kept."""
    text = ("1,2,3\n"
            "\n"
            " 4.5 , -6e2 ,7.25 \r\n"
            "x,1,2\n"          # header-like: dropped
            "1,2\n"            # two fields: dropped
            "8,9,10,11,12\n"   # extra fields ignored
            "1,,3\n"           # empty field: dropped
            "0x10,1,2\n"       # hex is not a decimal number: dropped
            "+1.,.5,-0\n"
            "inf,-Infinity,NaN\n"
            "1e40,1e-50,3\n"   # overflow to inf, underflow to 0
            "13,14,15")        # no final newline
    pts = gprt_hip.read_points(_write(tmp_path, "p.csv", text))
    want = np.array([[1, 2, 3], [4.5, -600, 7.25], [8, 9, 10], [1, 0.5, -0.0], [np.inf, -np.inf, np.nan],
                     [np.inf, 0, 3], [13, 14, 15]], np.float32)
    assert pts.dtype == np.float32 and pts.shape == want.shape
    np.testing.assert_array_equal(pts, want)
    assert np.signbit(pts[3, 2])


def test_read_points_unicode_trim_and_long_fields(tmp_path):
    """Fields are trimmed as Rust's str::trim does (Unicode White_Space, not
    only ASCII), and a number may be written with any number of digits."""
    nbsp, em, ideo, nel = "\u00a0", "\u2003", "\u3000", "\u0085"
    zeros = "0" * 5000
    text = (f"{nbsp}1,{em}2{ideo},3{nel}\n"
            f"\u2028{ideo}\n"                      # only White_Space: one empty field, dropped
            f"4,5\u200b,6\n"                       # U+200B is not White_Space: dropped
            f"0.{zeros}1,1{zeros}e-5000,-{zeros}7\n"
            f"{em}8,\u1680 9 \u205f,10\u202f\n")
    pts = gprt_hip.read_points(_write(tmp_path, "u.csv", text))
    want = np.array([[1, 2, 3], [np.float32("1e-5001"), 1, -7], [8, 9, 10]], np.float32)
    np.testing.assert_array_equal(pts, want)
    b = gprt_hip.read_points(_write(tmp_path, "u.txt", f"{ideo}{nbsp}\n1,2,3,{em}4{nbsp}\n"), gprt_hip.CSV_BODIES)
    np.testing.assert_array_equal(b, [[1, 2, 3, 4]])


def test_read_points_not_utf8(tmp_path):
    path = _write(tmp_path, "p.csv", b"1,2,3\n4,5,6,\xff\xfe\n7,8,9\n")
    np.testing.assert_array_equal(gprt_hip.read_points(path), [[1, 2, 3], [7, 8, 9]])
    with pytest.raises(gprt_hip.GprtError) as e:
        gprt_hip.read_points(_write(tmp_path, "b.txt", b"1,2,3,4\n\xc3(,1,1,1\n"), gprt_hip.CSV_BODIES)
    assert e.value.code == gprt_hip.ERR_INPUT


def test_read_points_exact_binary32(tmp_path):
    """Shortest round-trip decimal forms read back to the same binary32 values,
    in order, over a file large enough to be parsed in several slices."""
    rng = np.random.default_rng(5)
    v = np.concatenate([rng.normal(size=(150_000, 3)) * 1e3, rng.random((150_000, 3)) * 1e-3]).astype(np.float32)
    rows = [",".join(np.format_float_positional(x, unique=True, trim="-") if i % 2 else
                     np.format_float_scientific(x, unique=True, trim="-") for i, x in enumerate(r)) for r in v]
    path = _write(tmp_path, "big.csv", "\n".join(rows) + "\n")
    got = gprt_hip.read_points(path)
    assert got.shape == v.shape
    assert np.array_equal(got.view(np.uint32), v.view(np.uint32))


def test_read_bodies_rules(tmp_path):
    """barneshut.rs:14-31: blank lines and lines with < 4 fields skipped, bad
    fields read as 0, mass <= 0 skipped (the 0,0,0,0 dummy line)."""
    text = ("0,0,0,0\n"
            "1,2,3,4\n"
            "   \n"
            "a,2,3,5\n"       # bad x -> 0
            "1,2,3\n"         # three fields: skipped
            "1,2,3,-1\n"      # negative mass: skipped
            "1,2,3,junk\n"    # mass 0: skipped
            "7,8,9,0.5,extra\r\n")
    b = gprt_hip.read_points(_write(tmp_path, "g.txt", text), gprt_hip.CSV_BODIES)
    np.testing.assert_array_equal(b, np.array([[1, 2, 3, 4], [0, 2, 3, 5], [7, 8, 9, 0.5]], np.float32))


def test_read_points_empty_and_missing(tmp_path):
    assert gprt_hip.read_points(_write(tmp_path, "e.csv", "")).shape == (0, 3)
    assert gprt_hip.read_points(_write(tmp_path, "h.csv", "x,y,z\n")).shape == (0, 3)
    with pytest.raises(OSError):
        gprt_hip.read_points(str(tmp_path / "none.csv"))


def test_read_points_capacity(tmp_path):
    """Too small a buffer: nothing written, the row count returned, ERR_ARG."""
    import ctypes
    path = _write(tmp_path, "p.csv", "1,2,3\n4,5,6\n7,8,9\n")
    rows = ctypes.c_int64()
    rc = gprt_hip.lib().gprt_hip_read_points(path.encode(), gprt_hip.CSV_POINTS, None, 0, ctypes.byref(rows))
    assert rc == gprt_hip.ERR_ARG and rows.value == 3
    buf = np.full((2, 3), -1, np.float32)
    rc = gprt_hip.lib().gprt_hip_read_points(path.encode(), gprt_hip.CSV_POINTS, buf.ctypes.data, 2,
                                             ctypes.byref(rows))
    assert rc == gprt_hip.ERR_ARG and rows.value == 3 and (buf == -1).all()


# ---------------------------------------------------------------- harness logic
def test_parse_args():
    assert sandbox.parse_args(["s", "d.csv", "q.csv", "5"]) == ("d.csv", "q.csv", 5, pytest.approx(0.10), 3.0,
                                                                 False, False)
    d, q, k, p, m, dump, tune = sandbox.parse_args(["s", "d", "q", "8", "dump", "0.5", "oops", "2", "7", "tune"])
    assert (k, p, m, dump, tune) == (8, 0.5, 2.0, True, True)
    with pytest.raises(SystemExit):
        sandbox.parse_args(["s", "d", "q"])
    with pytest.raises(SystemExit):
        sandbox.parse_args(["s", "d", "q", "five"])


def test_choose_schedule(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    lines = []
    cfg, leaf, sah = sandbox.choose_schedule(None, None, None, 5, 0.25, 2.0, False, lines.append)
    assert lines == ["[LOAD] Using CLI defaults: P=0.25, M=2.0"]
    assert cfg.radius_param == 0.25 and cfg.radius_mult == 2.0 and cfg.sort_queries == 0 and leaf == 8
    assert sah is False
    # a reference wisdom file (no build_strategy / geom_type) loads
    (tmp_path / "gprt_wisdom.json").write_text('{"radius_increment_mult": 3.0, "max_hits_per_query": 2000, '
                                               '"use_morton_lbv": true, "radius_heuristic": '
                                               '{"SampledPercentile": 0.5}, "memory_strategy": '
                                               '"PayloadRegisterHeap"}')
    lines.clear()
    cfg, leaf, sah = sandbox.choose_schedule(None, None, None, 5, 0.25, 2.0, False, lines.append)
    assert lines == ["[LOAD] Loaded auto-tuned schedule from gprt_wisdom.json"]
    assert cfg.radius_param == 0.5 and cfg.sort_queries == 1
    assert sah is True  # build_strategy absent: Schedule::default(), PreferFastTrace
    # an unreadable file falls back to the command line without a message
    (tmp_path / "gprt_wisdom.json").write_text("not json")
    lines.clear()
    cfg, _, _ = sandbox.choose_schedule(None, None, None, 5, 0.25, 2.0, False, lines.append)
    assert lines == [] and cfg.radius_param == 0.25 and cfg.radius_mult == 2.0


def test_format_results():
    ids = np.array([[3, 1, 2], [7, gprt_hip.ID_NONE, gprt_hip.ID_NONE]], np.uint32)
    assert sandbox.format_results(ids) == "Q0:[3, 1, 2]\nQ1:[7]\n"
    assert sandbox.format_results(np.zeros((0, 3), np.uint32)) == ""


def test_rust_exponent_format():
    assert barneshut.rust_exp(1.23456e-3) == "1.2346e-3"
    assert barneshut.rust_exp(5.0) == "5.0000e0"
    assert barneshut.rust_exp(-2.5e11) == "-2.5000e11"
    assert barneshut.rust_exp(0.0) == "0.0000e0"
