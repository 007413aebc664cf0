"""Butteraugli on the GPU (libvship.so, vship_amd/csrc/butter.hip) against the
FP64 oracle (oracle/butter_oracle.c) and the reference's own output
(tests/golden/reference_butter.json), through the C ABI:

* the batched entry points (Vship_ButteraugliBatch*): norms within 1e-4
  relative of the oracle, the distortion map within 1e-4 relative + 1e-6
  absolute, on float and 16-bit inputs, odd sizes, 1x1 and 1920x1080, sizes
  on and just past the kernels' tile boundaries, Q-norms 1..5 and several
  intensities;
* the reference's per-frame entry points (Vship_ComputeButteraugli*): the int
  score fields are the norms truncated, the map written with the caller's
  stride;
* against the reference's FP32 output on the golden cases;
* identical images give zero; a pair's norms do not depend on the batch it is
  in (bit-exact).
"""
import json
import os

import numpy as np
import pytest

import vship_amd
from vship_amd import api
from tests import butter_lib as bl
from tests.golden.make_golden import BUTTER_CASES, butter_inputs

pytestmark = pytest.mark.gpu

NORM_RTOL = 1e-4
MAP_RTOL, MAP_ATOL = 1e-4, 1e-6
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_butter.json")


@pytest.fixture(scope="module", autouse=True)
def gpu_ready():
    vship_amd.gpu_full_check(0)
    yield


def say(*a):
    print(*a, flush=True)


def gpu_norms(pairs, qnorm=2, intensity=203.0, distmaps=False):
    h, w = np.asarray(pairs[0][0][0]).shape
    dp = api.ButterDevicePairs(pairs, distmaps=distmaps)
    try:
        with api.ButteraugliBatch(w, h, len(pairs), qnorm, intensity) as b:
            norms = b.compute(dp)
            maps = [dp.distmap(i) for i in range(len(pairs))] if distmaps else None
    finally:
        dp.free()
    return (norms, maps) if distmaps else norms


def check_norms(got, want, what):
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-12)
    say(f"{what}: gpu {got} oracle {want} max rel {rel.max():.2e}")
    assert np.all(rel <= NORM_RTOL), what


CASES = [
    ("f32", 64, 48, 2, 203.0, 21),
    ("u16", 75, 47, 2, 203.0, 22),
    ("f32", 97, 61, 3, 203.0, 23),
    ("u16", 33, 65, 1, 80.0, 24),
    ("f32", 7, 5, 2, 203.0, 25),
    ("f32", 130, 17, 5, 400.0, 26),
    ("u16", 256, 256, 2, 203.0, 27),
]


@pytest.mark.parametrize("fmt,w,h,q,inten,seed", CASES)
def test_batch_matches_oracle(fmt, w, h, q, inten, seed):
    say(f"case butter {fmt} {w}x{h} q{q} intensity {inten}")
    ref, dist = butter_inputs(w, h, fmt, seed)
    norms, maps = gpu_norms([(ref, dist)], q, inten, distmaps=True)
    want, wmap = bl.butteraugli(ref, dist, q, inten, diffmap=True)
    check_norms(norms[0], want, f"{fmt} {w}x{h}")
    err = np.abs(maps[0] - wmap) - (MAP_ATOL + MAP_RTOL * np.abs(wmap))
    say(f"map max abs diff {np.abs(maps[0] - wmap).max():.2e} of max {np.abs(wmap).max():.3f}")
    assert err.max() <= 0.0


# Sizes on and just past the kernels' tiles (blurs 64 x 32, Malta 64 x 16,
# mask 32 x 32) at full scale and, halved, at half scale: regions that end
# exactly at an image edge, cross it by one pixel, or are narrower than the
# widest blur window (21 pixels each side).
TILE_EDGE_SIZES = [(64, 32), (65, 33), (127, 31), (128, 64), (129, 17), (63, 97), (20, 130)]


@pytest.mark.parametrize("w,h", TILE_EDGE_SIZES)
def test_tile_edges_match_oracle(w, h):
    say(f"case butter tile edges {w}x{h}")
    ref, dist = butter_inputs(w, h, "f32", 100 + w + h)
    norms, maps = gpu_norms([(ref, dist)], distmaps=True)
    want, wmap = bl.butteraugli(ref, dist, 2, 203.0, diffmap=True)
    check_norms(norms[0], want, f"{w}x{h}")
    err = np.abs(maps[0] - wmap) - (MAP_ATOL + MAP_RTOL * np.abs(wmap))
    say(f"map max abs diff {np.abs(maps[0] - wmap).max():.2e} of max {np.abs(wmap).max():.3f}")
    assert err.max() <= 0.0


# 4:2:0 with left-located chroma takes the staged input kernel (8-bit and
# 16-bit containers); every other layout and location the generic one
YUV_CASES = [("420", 8, api.CHROMA_LEFT, api.MATRIX_BT709), ("420", 10, api.CHROMA_LEFT, api.MATRIX_BT709),
             ("420", 12, api.CHROMA_LEFT, api.MATRIX_BT2020NCL),
             ("422", 10, api.CHROMA_CENTER, api.MATRIX_BT2020NCL),
             ("444", 8, api.CHROMA_LEFT, api.MATRIX_BT601), ("410", 12, api.CHROMA_TOPLEFT, api.MATRIX_BT709)]


@pytest.mark.parametrize("layout,bits,loc,matrix", YUV_CASES)
def test_yuv_input_matches_oracle(layout, bits, loc, matrix):
    """YUV pairs: the input kernel's R'G'B' is the SSIMULACRA2 front end's
    (the oracle's so_yuv_to_rgb_f64), the rest as for RGB."""
    from tests.test_gpu_yuv_layouts import yuv_pair
    from tests import oracle_lib as ol

    f8, f16, sw, sh = api.YUV_LAYOUTS[layout]
    w, h = 133, 77
    say(f"case butter YUV{layout} {bits}-bit loc {loc} matrix {matrix} {w}x{h}")
    ref, dist = yuv_pair(w, h, layout, bits, 900 + bits)
    dp = api.ButterDevicePairs([(ref, dist)], fmt=f8 if bits == 8 else f16, bits=bits, matrix=matrix,
                               chroma_location=loc)
    try:
        with api.ButteraugliBatch(w, h, 1) as b:
            got = b.compute(dp)[0]
    finally:
        dp.free()
    rgb = [ol.yuv_to_rgb_any(im, bits, matrix, sw, sh, loc).astype(np.float32) for im in (ref, dist)]
    check_norms(got, bl.butteraugli(list(rgb[0]), list(rgb[1])), f"YUV{layout}")


def test_yuv420_8bit_ignores_bits_field():
    """For the 8-bit YUV formats the descriptor's bits field is ignored (the
    header says so): bits = 0 scores the same as bits = 8."""
    from tests.test_gpu_yuv_layouts import yuv_pair

    w, h = 96, 64
    say(f"case butter YUV420 8-bit bits=0 {w}x{h}")
    ref, dist = yuv_pair(w, h, "420", 8, 77)
    out = []
    for bits in (8, 0):
        dp = api.ButterDevicePairs([(ref, dist)], fmt=api.PF_YUV420_U8, bits=bits)
        try:
            with api.ButteraugliBatch(w, h, 1) as b:
                out.append(b.compute(dp)[0])
        finally:
            dp.free()
    assert np.array_equal(out[0], out[1])


def test_host_distmap_is_rejected():
    """A distance map in host memory is refused with BadPointer before any
    kernel runs (the reference's per-frame ABI writes its map to host memory,
    the batched ABI takes a device buffer)."""
    say("case butter host distmap rejected")
    ref, dist = butter_inputs(64, 48, "f32", 61)
    dp = api.ButterDevicePairs([(ref, dist)])
    host_map = np.zeros((48, 64), np.float32)
    dp.descs[0].distmap = host_map.ctypes.data
    dp.descs[0].distmap_stride = 64 * 4
    norms = np.zeros((1, 3))
    try:
        with api.ButteraugliBatch(64, 48, 1) as b:
            rc = api.lib().Vship_ButteraugliBatchCompute(b._h, dp.format, dp.descs, 1, norms.ctypes.data)
            assert rc == 10  # BadPointer
            # a device map whose stride is shorter than a row is refused too
            dp.descs[0].distmap = dp.buffer.ptr
            dp.descs[0].distmap_stride = 60 * 4
            rc = api.lib().Vship_ButteraugliBatchCompute(b._h, dp.format, dp.descs, 1, norms.ctypes.data)
            assert rc == 10
            # and the handle still works afterwards
            dp.descs[0].distmap = None
            got = b.compute(dp)[0]
    finally:
        dp.free()
    check_norms(got, bl.butteraugli(ref, dist), "after rejected maps")


@pytest.mark.parametrize("dtype", ["u8", "f16"])
def test_rgb_u8_f16_inputs_match_oracle(dtype):
    """Planar RGB u8 (v / 255) and f16 inputs against the oracle on the same
    values (u8 as the u16 code v * 257, f16 widened to f32, both exact)."""
    w, h = 83, 57
    say(f"case butter RGB {dtype} {w}x{h}")
    ref, dist = butter_inputs(w, h, "f32", 71)
    if dtype == "u8":
        gr = [np.round(p * 255).astype(np.uint8) for p in ref]
        gd = [np.round(p * 255).astype(np.uint8) for p in dist]
        fmt = api.PF_RGB_U8
        want = bl.butteraugli([p.astype(np.uint16) * 257 for p in gr], [p.astype(np.uint16) * 257 for p in gd])
    else:
        gr = [p.astype(np.float16) for p in ref]
        gd = [p.astype(np.float16) for p in dist]
        fmt = api.PF_RGB_F16
        want = bl.butteraugli([p.astype(np.float32) for p in gr], [p.astype(np.float32) for p in gd])
    dp = api.ButterDevicePairs([(gr, gd)], fmt=fmt)
    try:
        with api.ButteraugliBatch(w, h, 1) as b:
            got = b.compute(dp)[0]
    finally:
        dp.free()
    check_norms(got, want, f"RGB {dtype}")


def test_one_pixel():
    say("case butter 1x1")
    ref = [np.full((1, 1), v, np.float32) for v in (0.2, 0.5, 0.7)]
    dist = [np.full((1, 1), v, np.float32) for v in (0.25, 0.45, 0.7)]
    check_norms(gpu_norms([(ref, dist)])[0], bl.butteraugli(ref, dist), "1x1")


def test_1080p_matches_oracle():
    say("case butter 1920x1080 u16")
    ref, dist = butter_inputs(1920, 1080, "u16", 31)
    check_norms(gpu_norms([(ref, dist)])[0], bl.butteraugli(ref, dist), "1920x1080")


def test_identical_is_zero():
    say("case butter identical 96x80")
    ref, _ = butter_inputs(96, 80, "f32", 5)
    norms, maps = gpu_norms([(ref, [p.copy() for p in ref])], distmaps=True)
    assert np.all(norms == 0.0)
    assert np.all(maps[0] == 0.0)


def test_batch_independent_of_neighbours():
    say("case butter batch of 3 vs alone, 120x70")
    pairs = [butter_inputs(120, 70, "f32", s) for s in (41, 42, 43)]
    together = gpu_norms(pairs)
    for i, p in enumerate(pairs):
        assert np.array_equal(gpu_norms([p])[0], together[i])


def test_reference_abi_truncates_and_writes_map():
    say("case butter reference ABI 80x60")
    ref, dist = butter_inputs(80, 60, "f32", 51)
    batch = gpu_norms([(ref, dist)])[0]
    with api.ButteraugliHandler(80, 60) as h:
        ints, dm = h.compute_float(ref, dist, distmap=True)
        ints16 = h.compute_uint16([np.round(p * 65535).astype(np.uint16) for p in ref],
                                  [np.round(p * 65535).astype(np.uint16) for p in dist])
    assert ints == tuple(int(np.float32(v)) for v in batch)
    want, wmap = bl.butteraugli(ref, dist, 2, 203.0, diffmap=True)
    assert np.all(np.abs(dm - wmap) <= MAP_ATOL + MAP_RTOL * np.abs(wmap))
    u16 = [[np.round(p * 65535).astype(np.uint16) for p in im] for im in (ref, dist)]
    assert ints16 == tuple(int(np.float32(v)) for v in bl.butteraugli(*u16))


def test_reference_abi_errors():
    say("case butter reference ABI error codes")
    L = api.lib()
    bad = api.Handler(id=12345)
    score = api.ButterScore()
    planes = [np.zeros((4, 4), np.float32) for _ in range(3)]
    ptrs = api._plane_ptrs(planes)
    assert L.Vship_ComputeButteraugliFloat(bad, score, None, 0, ptrs, ptrs, 16) == 9  # BadHandler
    assert L.Vship_ButteraugliFree(bad) == 9
    with api.ButteraugliHandler(4, 4) as h:
        # stride shorter than a row
        assert L.Vship_ComputeButteraugliFloat(h._h, score, None, 0, ptrs, ptrs, 8) == 10  # BadPointer


@pytest.mark.parametrize("case", [c[0] for c in BUTTER_CASES])
def test_matches_reference_hip_output(case):
    """GPU against the reference's own FP32 HIP code on the golden inputs."""
    say(f"case butter golden {case}")
    with open(GOLDEN) as fh:
        g = {c["name"]: c for c in json.load(fh)["cases"]}[case]
    name, w, h, fmt, seed, q = next(c for c in BUTTER_CASES if c[0] == case)
    ref, dist = butter_inputs(w, h, fmt, seed)
    norms, maps = gpu_norms([(ref, dist)], q, 203.0, distmaps=True)
    check_norms(norms[0], np.array(g["norms"]), f"golden {case}")
    got = np.array([maps[0][y, x] for x, y, _ in g["map_samples"]])
    want = np.array([v for _, _, v in g["map_samples"]])
    assert np.all(np.abs(got - want) <= 1e-5 + 1e-3 * np.abs(want))
