"""The FFVship-style command line (vship_amd/ffvship, vship_amd/csrc/ffvship_cli.cpp)
on the CPU: YUV4MPEG2 header parsing, the reference's frame selection
(src/FFVship.cpp:262-330, with the worked examples of doc/FFVship.md), the
JSON and statistics formats (src/FFVship.cpp:142-200,439-468) of both
metrics, and the
errors the command reports before it touches a GPU."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from vship_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOOKS = os.path.join(ROOT, "tests", "_build", "libffvship_host.so")
CLI = os.path.join(ROOT, "vship_amd", "ffvship")

needs_hooks = pytest.mark.skipif(not os.path.exists(HOOKS), reason="test helpers not built (python -m vship_amd.build)")
needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="command line not built (python -m vship_amd.build)")


def hooks():
    L = ctypes.CDLL(HOOKS)
    L.ffv_parse_header.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]
    I = ctypes.POINTER(ctypes.c_int)
    L.ffv_select.argtypes = [ctypes.c_int] * 6 + [ctypes.c_char_p, ctypes.c_char_p, I, I, ctypes.c_int,
                                                  ctypes.c_char_p, ctypes.c_int]
    F = ctypes.POINTER(ctypes.c_float)
    L.ffv_stats.argtypes = [F, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    L.ffv_json.argtypes = [F, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    L.ffv_json_norms.argtypes = [F, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    L.ffv_metric.argtypes = [ctypes.c_char_p]
    return L


def header(line):
    out = (ctypes.c_int * 6)()
    err = ctypes.create_string_buffer(512)
    rc = hooks().ffv_parse_header(line.encode(), out, err, len(err))
    return (rc, list(out), err.value.decode())


def select(ts, te, start=0, end=-1, every=1, offset=0, sidx="", eidx=""):
    a, b = (ctypes.c_int * 4096)(), (ctypes.c_int * 4096)()
    err = ctypes.create_string_buffer(512)
    n = hooks().ffv_select(ts, te, start, end, every, offset, sidx.encode(), eidx.encode(), a, b, 4096, err, len(err))
    if n < 0:
        return None, err.value.decode()
    return (list(a[:n]), list(b[:n])), ""


@needs_hooks
@pytest.mark.parametrize("tag,fmt,bits,loc", [
    ("C420jpeg", api.PF_YUV420_U8, 8, api.CHROMA_CENTER), ("", api.PF_YUV420_U8, 8, api.CHROMA_CENTER),
    ("C420", api.PF_YUV420_U8, 8, api.CHROMA_CENTER), ("C420mpeg2", api.PF_YUV420_U8, 8, api.CHROMA_LEFT),
    ("C420paldv", api.PF_YUV420_U8, 8, api.CHROMA_TOPLEFT), ("C422", api.PF_YUV422_U8, 8, api.CHROMA_LEFT),
    ("C444", api.PF_YUV444_U8, 8, api.CHROMA_LEFT), ("C411", api.PF_YUV411_U8, 8, api.CHROMA_LEFT),
    ("C420p10", api.PF_YUV420_U16, 10, api.CHROMA_LEFT), ("C422p12", api.PF_YUV422_U16, 12, api.CHROMA_LEFT),
    ("C444p16", api.PF_YUV444_U16, 16, api.CHROMA_LEFT)])
def test_y4m_colorspaces(tag, fmt, bits, loc):
    w, h = 33, 17
    rc, out, err = header(f"YUV4MPEG2 W{w} H{h} F25:1 Ip A1:1 {tag} XYSCSS=ANY".replace("  ", " "))
    assert rc == 0, err
    _, _, sw, sh = next(v for v in api.YUV_LAYOUTS.values() if fmt in v[:2])
    bps = 2 if bits > 8 else 1
    cw, ch = -(-w // (1 << sw)), -(-h // (1 << sh))
    assert out == [w, h, fmt, bits, loc, bps * (w * h + 2 * cw * ch)]


@needs_hooks
@pytest.mark.parametrize("line,msg", [
    ("YUV4MPEG2 W64 H32 C420 XCOLORRANGE=FULL", "full-range"), ("YUV4MPEG2 W64 H32 Cmono", "unsupported"),
    ("YUV4MPEG2 W64 H32 C444alpha", "unsupported"), ("YUV4MPEG2 H32 C420", "W/H"), ("RIFF W64 H32", "YUV4MPEG2"),
    ("YUV4MPEG2 W64 H32 C420p7", "unsupported")])
def test_y4m_rejects(line, msg):
    rc, _, err = header(line)
    assert rc == 1 and msg in err


@needs_hooks
def test_selection_matches_the_reference_examples():
    # doc/FFVship.md "encoded-offset": --start 0 --end 5 --encoded-offset 1 / -1 on five-frame clips
    assert select(5, 5, 0, 5, 1, 1)[0] == ([0, 1, 2, 3], [1, 2, 3, 4])
    assert select(5, 5, 0, 4, 1, 1)[0] == ([0, 1, 2, 3], [1, 2, 3, 4])
    assert select(5, 5, 0, 5, 1, -1)[0] == ([1, 2, 3, 4], [0, 1, 2, 3])
    assert select(5, 5, 1, 5, 1, -1)[0] == ([1, 2, 3, 4], [0, 1, 2, 3])
    # doc/FFVship.md "--source-indices": --start 10 --every 2 --encoded-offset 1 --source-indices 13
    assert select(100, 100, 10, -1, 2, 1, "13")[0] == ([36], [37])


@needs_hooks
def test_selection_rules():
    assert select(10, 10)[0] == (list(range(10)), list(range(10)))
    assert select(10, 10, 2, 9, 3)[0] == ([2, 5, 8], [2, 5, 8])
    assert select(10, 6)[0] == (list(range(6)), list(range(6)))           # the shorter clip bounds the range
    assert select(10, 10, -5, 99)[0] == (list(range(10)), list(range(10)))  # start and end are clamped
    assert select(10, 10, 7, 3)[0] == ([], [])
    assert select(10, 10, 0, -1, 1, 0, "1,3", "0,2")[0] == ([1, 3], [0, 2])
    assert select(10, 10, 0, -1, 1, 2, "4")[0] == ([4], [6])
    assert select(10, 10, 0, -1, 1, 20)[1].startswith("encoded_offset 20 does not allow")
    assert select(10, 10, 0, -1, 1, 0, "10")[1].startswith("Source Invalid frame index found 10")
    assert select(10, 10, 0, -1, 1, 3, "8")[1].startswith("Encoded Invalid frame index found 8")
    assert select(10, 10, 0, -1, 1, 0, "1,2", "1")[1].startswith("Source indices and Encoded indices are of different")
    assert select(10, 10, 0, -1, 0)[1].startswith("--every")


def stats(data, label="SSIMULACRA2"):
    a = np.ascontiguousarray(data, np.float32)
    buf = ctypes.create_string_buffer(4096)
    hooks().ffv_stats(a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(a), label.encode(), buf, len(buf))
    return buf.value.decode()


@needs_hooks
def test_statistics_block():
    data = np.arange(1, 41, dtype=np.float32)[::-1]  # 40 values, unsorted
    lines = stats(data).split("\n")
    # label centred in a rule as wide as the 18-character name + " : " + 12-character value
    assert lines[0] == "-" * 11 + "SSIMULACRA2" + "-" * 11
    v = np.sort(data)
    mean = v.mean(dtype=np.float64)
    expect = [("Average", mean), ("Standard Deviation", np.sqrt(((v - mean) ** 2).mean())), ("Median", v[20]),
              ("5th percentile", v[2]), ("95th percentile", v[38]), ("Minimum", v[0]), ("Maximum", v[-1])]
    for line, (name, val) in zip(lines[1:8], expect):
        assert line == f"{name:>18} : {val:>12.6f}"
    assert lines[8] == "" and stats([]) == ""


@needs_hooks
def test_json_format():
    a = np.array([61.5, 100.0, -3.25], np.float32)
    buf = ctypes.create_string_buffer(256)
    hooks().ffv_json(a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 3, buf, len(buf))
    assert buf.value.decode() == "[[61.5], [100], [-3.25]]"


@needs_hooks
def test_butteraugli_json_format_and_metric_names():
    # src/FFVship.cpp:448-452: [[q, 3, inf], ...] with ", " between values
    a = np.array([2.5, 2.75, 4.0, 1.0, 1.25, 3.5], np.float32)
    buf = ctypes.create_string_buffer(256)
    hooks().ffv_json_norms(a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 2, buf, len(buf))
    assert buf.value.decode() == "[[2.5, 2.75, 4], [1, 1.25, 3.5]]"
    # ffmpegmain.hpp:770-780: case-insensitive names and their short forms
    names = {"SSIMULACRA2": 0, "ssimu2": 0, "Butteraugli": 1, "BUTTER": 1, "cvvdp": 2, "psnr": -1}
    for n, v in names.items():
        assert hooks().ffv_metric(n.encode()) == v


def write_y4m(path, frames, w, h, tag="C420mpeg2", extra=""):
    with open(path, "wb") as f:
        f.write(f"YUV4MPEG2 W{w} H{h} F25:1 Ip A1:1 {tag}{extra}\n".encode())
        for planes in frames:
            f.write(b"FRAME\n")
            for p in planes:
                f.write(np.ascontiguousarray(p).tobytes())


def run_cli(*args):
    r = subprocess.run([CLI, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    return r.returncode, r.stdout, r.stderr


@needs_cli
def test_cli_errors_before_the_gpu(tmp_path):
    w, h = 32, 16
    fr = [np.zeros((h, w), np.uint8), np.zeros((h // 2, w // 2), np.uint8), np.zeros((h // 2, w // 2), np.uint8)]
    a, b, c = str(tmp_path / "a.y4m"), str(tmp_path / "b.y4m"), str(tmp_path / "c.y4m")
    write_y4m(a, [fr, fr], w, h)
    write_y4m(b, [fr], w, h, extra=" XCOLORRANGE=FULL")
    fr444 = [np.zeros((h, w), np.uint8)] * 3
    write_y4m(c, [fr444], w, h, tag="C444")
    assert run_cli("--source", a, "--encoded", str(tmp_path / "missing.y4m"))[2].startswith("cannot open")
    assert "full-range" in run_cli("--source", a, "--encoded", b)[2]
    assert "differ in size, layout or depth" in run_cli("--source", a, "--encoded", c)[2]
    assert "not supported by this build" in run_cli("-s", a, "-e", a, "-m", "CVVDP")[2]
    assert "not supported by this build" in run_cli("-s", a, "-e", a, "-m", "psnr")[2]
    assert "does not allow comparing" in run_cli("-s", a, "-e", a, "--encoded-offset", "5")[2]
    assert "unknown option" in run_cli("-s", a, "-e", a, "--frobnicate")[2]
    rc, out, _ = run_cli("--version")
    assert rc == 0 and out.startswith("FFVship 3.1.0-b")
    # a truncated last frame is dropped, a malformed frame header is an error
    with open(a, "ab") as f:
        f.write(b"FRAME\n" + b"\x00" * 10)
    assert "does not allow comparing" in run_cli("-s", a, "-e", a, "--encoded-offset", "3")[2]  # 2 frames, not 3
    # an empty selection needs no GPU
    rc, out, _ = run_cli("-s", a, "-e", a, "--encoded-offset", "2", "--live-score-output")
    assert rc == 0 and out == "0\n"
    with open(c, "ab") as f:
        f.write(b"FRAMX\n")
    assert "malformed FRAME header" in run_cli("-s", c, "-e", c)[2]
