"""Parity of the HIP path with the FP64 oracle, through the C ABI.

Tolerance (FP32 compute): |score_gpu - score_oracle| <= 1e-3, and each of
the 108 features within 2e-5 absolute + 1e-3 relative. FP16 planes: the
measured error is printed and bounded loosely (it is a storage-precision
variant). Exactness checks (identical inputs -> 100, batch composition) are
bit-exact.
"""
import ctypes
import threading

import numpy as np
import pytest

import vship_amd
from vship_amd import api, synth
from tests import oracle_lib as ol
from tests.gpu_helpers import FMT_NAMES, gpu_batch, oracle_rgb, rgb_pair, say

pytestmark = pytest.mark.gpu

SCORE_TOL = 1e-3


def feat_close(g, o):
    return np.all(np.abs(g - o) <= 2e-5 + 1e-3 * np.abs(o))


@pytest.fixture(scope="module", autouse=True)
def gpu_ready():
    say("gpu parity: full check of device 0")
    vship_amd.gpu_full_check(0)
    yield


def test_device_info():
    info = vship_amd.device_info(0)
    say(f"device: {info}")
    assert info["WarpSize"] == 64
    assert info["MultiProcessorCount"] >= 1
    assert vship_amd.device_count() >= 1


def test_config0_rgb8_matches_oracle():
    ref, dist = synth.config0()
    say("case: config0 256x256 RGB8, batch API")
    s_o, f_o = oracle_rgb(ref, dist, api.PF_RGB_U8)
    (s_g,), (f_g,) = gpu_batch([(ref, dist)], api.PF_RGB_U8)
    say(f"config0 oracle {s_o:.6f} gpu {s_g:.6f} diff {s_g - s_o:.2e}")
    assert abs(s_g - s_o) <= SCORE_TOL
    assert feat_close(f_g, f_o)


def test_config0_through_reference_entry_points():
    ref, dist = synth.config0()
    s_o, _ = oracle_rgb(ref, dist, api.PF_RGB_U8)
    say("case: config0 via Vship_ComputeSSIMU2Uint16v2 / Floatv2")
    with vship_amd.SSIMU2Handler(256, 256) as h:
        s16 = h.compute_uint16([p.astype(np.uint16) * 257 for p in ref], [p.astype(np.uint16) * 257 for p in dist])
        s32 = h.compute_float([p.astype(np.float32) / 255 for p in ref], [p.astype(np.float32) / 255 for p in dist])
    say(f"uint16 {s16:.6f} float {s32:.6f} oracle {s_o:.6f}")
    assert abs(s16 - s_o) <= SCORE_TOL
    assert abs(s32 - s_o) <= SCORE_TOL


def test_identical_padded_images_score_exactly_100():
    # the reference's own known answer (test/basic_metric_math_test.cpp:40-58):
    # zero images, one with 0xFF row padding, any size -> exactly 100
    w, h, stride = 1076, 240, 1100 * 4
    a = np.zeros((3, h, stride), np.uint8)
    b = np.zeros((3, h, stride), np.uint8)
    b[:, :, w * 4:] = 255
    say("case: identical zero images with padded rows, Vship_ComputeSSIMU2Float")
    L = api.lib()
    hd = api.Handler()
    assert L.Vship_SSIMU2Init(ctypes.byref(hd), w, h) == 0
    try:
        score = ctypes.c_double()
        pa = api.P3(*[a[i].ctypes.data for i in range(3)])
        pb = api.P3(*[b[i].ctypes.data for i in range(3)])
        assert L.Vship_ComputeSSIMU2Float(hd, ctypes.byref(score), pa, pb, stride) == 0
        assert score.value == 100.0
    finally:
        assert L.Vship_SSIMU2Free(hd) == 0


@pytest.mark.parametrize("fmt", [api.PF_RGB_U8, api.PF_RGB_U16, api.PF_RGB_F16, api.PF_RGB_F32])
def test_identical_random_images_score_exactly_100(fmt):
    ref, _ = rgb_pair(97, 61, fmt, 5)
    say(f"case: identical random {FMT_NAMES[fmt]} 97x61")
    (s,), (f,) = gpu_batch([(ref, [p.copy() for p in ref])], fmt)
    assert s == 100.0
    assert np.all(f == 0.0)


SIZES = [(1, 1), (3, 5), (17, 9), (33, 65), (70, 31), (129, 67), (200, 150)]


@pytest.mark.parametrize("fmt", [api.PF_RGB_U8, api.PF_RGB_U16, api.PF_RGB_F16, api.PF_RGB_F32])
def test_odd_sizes_all_formats(fmt):
    for i, (w, h) in enumerate(SIZES):
        ref, dist = rgb_pair(w, h, fmt, 100 + i)
        say(f"case: {FMT_NAMES[fmt]} {w}x{h}")
        s_o, f_o = oracle_rgb(ref, dist, fmt)
        (s_g,), (f_g,) = gpu_batch([(ref, dist)], fmt)
        assert abs(s_g - s_o) <= SCORE_TOL, (w, h, s_g, s_o)
        assert feat_close(f_g, f_o), (w, h, np.max(np.abs(f_g - f_o)))


@pytest.mark.parametrize("th", [48, 112])
def test_tile_heights_agree(th):
    ref, dist = rgb_pair(300, 211, api.PF_RGB_F32, 7)
    say(f"case: map tile height {th}")
    s_o, f_o = oracle_rgb(ref, dist, api.PF_RGB_F32)
    (s_g,), (f_g,) = gpu_batch([(ref, dist)], api.PF_RGB_F32, th=th)
    assert abs(s_g - s_o) <= SCORE_TOL
    assert feat_close(f_g, f_o)


def test_yuv420_8bit_small_and_odd():
    for (w, h, seed) in [(64, 48, 1), (97, 33, 2), (250, 141, 3)]:
        clip = synth.video_clip(w, h, 8, 2, seed)
        r, d = clip.frame(1)
        say(f"case: yuv420 8-bit BT.709 {w}x{h}")
        s_o, f_o = ol.score_yuv420(r, d, 8, ol.SO_MAT_BT709)
        (s_g,), (f_g,) = gpu_batch([(r, d)], api.PF_YUV420_U8)
        assert abs(s_g - s_o) <= SCORE_TOL, (w, h, s_g, s_o)
        assert feat_close(f_g, f_o)


def test_yuv420_1080p_config1_frames():
    clip = synth.video_clip(1920, 1080, 8, 240, 1)
    pairs = [clip.frame(f) for f in (0, 131)]
    say("case: config1 1080p YUV420P8 frames 0 and 131")
    s_g, f_g = gpu_batch(pairs, api.PF_YUV420_U8)
    for i, (r, d) in enumerate(pairs):
        s_o, f_o = ol.score_yuv420(r, d, 8, ol.SO_MAT_BT709)
        say(f"frame {i}: oracle {s_o:.6f} gpu {s_g[i]:.6f} diff {s_g[i] - s_o:.2e}")
        assert abs(s_g[i] - s_o) <= SCORE_TOL
        assert feat_close(f_g[i], f_o)


def test_yuv420_10bit_bt2020():
    clip = synth.video_clip(320, 180, 10, 3, 9)
    r, d = clip.frame(2)
    say("case: yuv420 10-bit BT.2020 320x180")
    s_o, f_o = ol.score_yuv420(r, d, 10, ol.SO_MAT_BT2020NCL)
    (s_g,), (f_g,) = gpu_batch([(r, d)], api.PF_YUV420_U16, bits=10, matrix=api.MATRIX_BT2020NCL)
    assert abs(s_g - s_o) <= SCORE_TOL
    assert feat_close(f_g, f_o)


def test_batch_composition_is_bit_exact():
    # each pair's result is the same alone and inside an uneven batch
    specs = [(640, 360, 1), (97, 61, 2), (1280, 720, 3), (33, 17, 4), (320, 240, 5)]
    pairs = []
    for w, h, seed in specs:
        c = synth.video_clip(w, h, 8, 1, seed)
        pairs.append(c.frame(0))
    say("case: uneven batch of 5 vs each alone")
    s_mix, f_mix = gpu_batch(pairs, api.PF_YUV420_U8)
    for i, p in enumerate(pairs):
        (s1,), (f1,) = gpu_batch([p], api.PF_YUV420_U8)
        assert s1 == s_mix[i], (i, specs[i], s1, s_mix[i])
        assert np.array_equal(f1, f_mix[i])


def test_persistent_grid_and_graph_replays_match_eager():
    specs = [[(200, 120, 1), (64, 64, 2)], [(333, 111, 3)], [(128, 96, 4), (90, 50, 5), (250, 250, 6)]]
    batches = []
    for bs in specs:
        pairs = []
        for w, h, seed in bs:
            pairs.append(rgb_pair(w, h, api.PF_RGB_U8, seed))
        batches.append(pairs)
    eager = [gpu_batch(p, api.PF_RGB_U8) for p in batches]
    say("case: persistent grid of 7 workgroups")
    for p, (s, f) in zip(batches, eager):
        s2, f2 = gpu_batch(p, api.PF_RGB_U8, persistent_grid=7)
        assert np.array_equal(s, s2) and np.array_equal(f, f2)
    say("case: one captured graph replayed over three batch shapes, twice")
    all_pairs = [x for p in batches for x in p]
    dp = [api.DevicePairs(p, api.PF_RGB_U8) for p in batches]
    with api.SSIMU2Batch(4, 333, 250, use_graph=True) as b:
        for rep in range(2):
            for k, d in enumerate(dp):
                s, f = b.compute(d.descs, features=True)
                assert np.array_equal(s, eager[k][0]), (rep, k)
                assert np.array_equal(f, eager[k][1])
                # eager work between replays
                gpu_batch(batches[k][:1], api.PF_RGB_U8)
    for d in dp:
        d.free()


def test_fp16_planes_error_is_small():
    # FP16 storage of scales 0 and 1 (FP32 for scales 2-5): the score moves by
    # a few 1e-4; bound 1e-3 against the FP64 oracle
    cases = [(640, 360, 8, 11), (960, 540, 10, 12), (200, 120, 10, 13), (1280, 720, 8, 14)]
    pairs, bits_l, expect = [], [], []
    for w, h, bits, seed in cases:
        clip = synth.video_clip(w, h, bits, 1, seed)
        r, d = clip.frame(0)
        m = ol.SO_MAT_BT709 if bits == 8 else ol.SO_MAT_BT2020NCL
        pairs.append((r, d))
        bits_l.append(bits)
        expect.append(ol.score_yuv420(r, d, bits, m)[0])
    # one uneven batch mixing 8- and 10-bit pairs
    dp = api.DevicePairs(pairs, api.PF_YUV420_U8, 8, api.MATRIX_BT709)
    for i, b in enumerate(bits_l):
        if b != 8:
            dp.descs[i].format, dp.descs[i].bits, dp.descs[i].matrix = api.PF_YUV420_U16, b, api.MATRIX_BT2020NCL
    with api.SSIMU2Batch(len(pairs), 1280, 720, fp16=True) as b16, api.SSIMU2Batch(len(pairs), 1280, 720) as b32:
        s16 = b16.compute(dp.descs)
        s32 = b32.compute(dp.descs)
    dp.free()
    for k, (w, h, bits, _) in enumerate(cases):
        say(f"case: fp16 planes {w}x{h} {bits}-bit: oracle {expect[k]:.6f} fp16 {s16[k]:.6f} "
            f"diff {s16[k] - expect[k]:.2e} (fp32 planes diff {s32[k] - expect[k]:.2e})")
    assert np.all(np.abs(s16 - np.array(expect)) <= 1e-3)
    assert np.all(np.abs(s32 - np.array(expect)) <= SCORE_TOL)


def test_several_handles_in_threads():
    pairs = [rgb_pair(160, 90, api.PF_RGB_F32, 40 + i) for i in range(3)]
    expect = [oracle_rgb(r, d, api.PF_RGB_F32)[0] for r, d in pairs]
    out = [None] * 3
    say("case: three handles used from three threads")

    def work(i):
        with vship_amd.SSIMU2Handler(160, 90) as h:
            out[i] = [h.compute_float(*pairs[i]) for _ in range(5)]

    ts = [threading.Thread(target=work, args=(i,)) for i in range(3)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    for i in range(3):
        assert len(set(out[i])) == 1
        assert abs(out[i][0] - expect[i]) <= SCORE_TOL


def test_error_paths():
    L = api.lib()
    assert L.Vship_SetDevice(10 ** 6) == 7
    assert L.Vship_GPUFullCheck(-1) == 7
    info = api.DeviceInfo()
    assert L.Vship_GetDeviceInfo(ctypes.byref(info), 999) == 7
    say("case: batch with a host pointer and with a short stride are rejected before any launch")
    ref, dist = rgb_pair(32, 16, api.PF_RGB_U8, 1)
    dp = api.DevicePairs([(ref, dist)], api.PF_RGB_U8)
    host = np.zeros(32 * 16, np.uint8)
    with api.SSIMU2Batch(1, 32, 16) as b:
        bad = (api.Pair * 1)()
        ctypes.memmove(ctypes.addressof(bad), ctypes.addressof(dp.descs), ctypes.sizeof(api.Pair))
        bad[0].ref[0] = host.ctypes.data
        with pytest.raises(api.VshipError) as e:
            b.compute(bad)
        assert e.value.code == 10
        bad2 = (api.Pair * 1)()
        ctypes.memmove(ctypes.addressof(bad2), ctypes.addressof(dp.descs), ctypes.sizeof(api.Pair))
        bad2[0].dist_stride[1] = 8
        with pytest.raises(api.VshipError) as e:
            b.compute(bad2)
        assert e.value.code == 10
        # too many pairs for the handle
        with pytest.raises(api.VshipError):
            b.compute((api.Pair * 2)(dp.descs[0], dp.descs[0]))
        s = b.compute(dp.descs)  # still usable
        assert 0 < s[0] <= 100
    dp.free()
