"""The matrix-core map kernel (map_tile_height = 1: tiles of 64 columns by
128 rows, both blur passes as 16x16x32 MFMA products of operands split into
16-bit parts: a scaled two-way fp16 split for the bounded 8- to 16-bit RGB and
YUV formats, an exact three-way bf16 split for float RGB) against the FP64 oracle
(|dscore| <= 1e-3, features as in the parity tests) and against the
vector-ALU kernel (the same FP32 maps to rounding: |dscore| <= 1e-4)."""
import numpy as np
import pytest

import vship_amd
from vship_amd import api, synth
from tests import oracle_lib as ol
from tests.gpu_helpers import FMT_NAMES, gpu_batch, oracle_rgb, rgb_pair, say

pytestmark = pytest.mark.gpu

MFMA = 1
SCORE_TOL = 1e-3


def feat_close(g, o):
    return np.all(np.abs(g - o) <= 2e-5 + 1e-3 * np.abs(o))


@pytest.fixture(scope="module", autouse=True)
def gpu_ready():
    vship_amd.gpu_full_check(0)
    yield


@pytest.mark.parametrize("fmt", [api.PF_RGB_U8, api.PF_RGB_U16, api.PF_RGB_F16, api.PF_RGB_F32])
def test_mfma_sizes_and_formats(fmt):
    # sizes with FP32-class segments (>= 32768 px) that are multiples of the
    # 64 x 128 tile, odd, thin, wider than tall, and with interior tiles above
    # a last tile row of 18 rows
    for i, (w, h) in enumerate([(256, 256), (300, 211), (193, 177), (1031, 33), (64, 600), (517, 389), (401, 530)]):
        ref, dist = rgb_pair(w, h, fmt, 300 + i)
        say(f"case: mfma map, {FMT_NAMES[fmt]} {w}x{h}")
        s_o, f_o = oracle_rgb(ref, dist, fmt)
        (s_m,), (f_m,) = gpu_batch([(ref, dist)], fmt, th=MFMA)
        (s_v,), (f_v,) = gpu_batch([(ref, dist)], fmt, th=48)
        say(f"  oracle {s_o:.7f} mfma {s_m:.7f} ({s_m - s_o:+.1e}) valu {s_v:.7f} ({s_v - s_o:+.1e})")
        assert abs(s_m - s_o) <= SCORE_TOL, (w, h, s_m, s_o)
        assert feat_close(f_m, f_o), (w, h, np.max(np.abs(f_m - f_o)))
        assert abs(s_m - s_v) <= 1e-4


def test_mfma_identical_inputs_score_exactly_100():
    ref, _ = rgb_pair(333, 250, api.PF_RGB_F32, 9)
    say("case: mfma map, identical 333x250 images")
    (s,), (f,) = gpu_batch([(ref, ref)], api.PF_RGB_F32, th=MFMA)
    assert s == 100.0
    assert np.all(f == 0)


def test_mfma_yuv_1080p_and_10bit():
    clip = synth.video_clip(1920, 1080, 8, 240, 1)
    pairs = [clip.frame(f) for f in (0, 77)]
    say("case: mfma map, 1920x1080 YUV420P8 frames 0 and 77")
    s_m = gpu_batch(pairs, api.PF_YUV420_U8, th=MFMA, features=False)
    for k, (r, d) in enumerate(pairs):
        s_o, _ = ol.score_yuv420(r, d, 8, ol.SO_MAT_BT709)
        say(f"  oracle {s_o:.7f} mfma {s_m[k]:.7f} ({s_m[k] - s_o:+.1e})")
        assert abs(s_m[k] - s_o) <= SCORE_TOL
    clip10 = synth.video_clip(960, 540, 10, 3, 4)
    r, d = clip10.frame(2)
    say("case: mfma map, 960x540 YUV420P10 BT.2020, FP32 and FP16 planes")
    s_o, _ = ol.score_yuv420(r, d, 10, ol.SO_MAT_BT2020NCL)
    (s32,) = gpu_batch([(r, d)], api.PF_YUV420_U16, bits=10, matrix=api.MATRIX_BT2020NCL, th=MFMA, features=False)
    (s16,) = gpu_batch([(r, d)], api.PF_YUV420_U16, bits=10, matrix=api.MATRIX_BT2020NCL, th=MFMA, fp16=True,
                       features=False)
    say(f"  oracle {s_o:.7f} mfma fp32 planes {s32:.7f} fp16 planes {s16:.7f}")
    assert abs(s32 - s_o) <= SCORE_TOL
    assert abs(s16 - s_o) <= SCORE_TOL


def test_mfma_batch_grid_and_graph_are_bit_exact():
    pairs = [rgb_pair(w, h, api.PF_RGB_U16, 40 + i) for i, (w, h) in enumerate([(400, 300), (190, 190), (640, 360)])]
    dp = api.DevicePairs(pairs, api.PF_RGB_U16)
    say("case: mfma map, uneven batch of 3: alone vs batch vs persistent grid vs graph replays")
    alone = []
    one = (api.Pair * 1)()
    with api.SSIMU2Batch(1, 640, 360, map_tile_height=MFMA) as b:
        for i in range(3):
            one[0] = dp.descs[i]
            alone.append(b.compute(one, 1)[0])
    alone = np.array(alone)
    with api.SSIMU2Batch(3, 640, 360, map_tile_height=MFMA) as b:
        assert np.array_equal(b.compute(dp.descs), alone)
    with api.SSIMU2Batch(3, 640, 360, map_tile_height=MFMA, persistent_grid=300) as b:
        assert np.array_equal(b.compute(dp.descs), alone)
    with api.SSIMU2Batch(3, 640, 360, map_tile_height=MFMA, use_graph=True) as b:
        for _ in range(3):
            assert np.array_equal(b.compute(dp.descs), alone)
    dp.free()


def test_mfma_extreme_yuv_codes_stay_in_range():
    # uniform noise over every 8-bit code in all three planes reaches the
    # extremes of the XYB range the scaled fp16 split is sized for
    rng = np.random.default_rng(77)
    w, h = 320, 240
    mk = lambda: [rng.integers(0, 256, (h, w)).astype(np.uint8),
                  rng.integers(0, 256, (h // 2, w // 2)).astype(np.uint8),
                  rng.integers(0, 256, (h // 2, w // 2)).astype(np.uint8)]
    pairs = [(mk(), mk()) for _ in range(2)]
    smooth = synth.video_clip(w, h, 8, 1, 3).frame(0)
    pairs.append((pairs[0][0], smooth[1]))  # extreme reference against a natural frame
    say("case: mfma map, 320x240 YUV420P8 with uniformly random codes (extreme XYB)")
    s_m = gpu_batch(pairs, api.PF_YUV420_U8, th=MFMA, features=False)
    s_v = gpu_batch(pairs, api.PF_YUV420_U8, th=48, features=False)
    for k, (r, d) in enumerate(pairs):
        s_o, _ = ol.score_yuv420(r, d, 8, ol.SO_MAT_BT709)
        say(f"  oracle {s_o:.6f} mfma {s_m[k]:.6f} ({s_m[k] - s_o:+.1e}) valu {s_v[k]:.6f}")
        assert np.isfinite(s_m[k])
        assert abs(s_m[k] - s_o) <= SCORE_TOL
        # the two GPU kernels agree to 1e-4 plus a few FP32 ulps of the
        # score (these noise frames score near -1163, where one ulp is 1.2e-4)
        assert abs(s_m[k] - s_v[k]) <= 1e-4 + 4 * 2.0 ** -23 * abs(s_o)


def test_mfma_yuv10_codes_above_depth_are_clamped():
    clip = synth.video_clip(320, 240, 10, 1, 6)
    r, d = clip.frame(0)
    wild = [p.copy() for p in d]
    wild[0][10:20, 10:20] = 65535
    wild[1][5:9, 5:9] = 3000
    clamped = [np.minimum(p, 1023) for p in wild]
    say("case: mfma map, 320x240 YUV420P10 with codes above 1023")
    s = gpu_batch([(r, wild), (r, clamped)], api.PF_YUV420_U16, bits=10, matrix=api.MATRIX_BT2020NCL, th=MFMA,
                  features=False)
    s_o, _ = ol.score_yuv420(r, wild, 10, ol.SO_MAT_BT2020NCL)
    say(f"  wild {s[0]:.6f} clamped {s[1]:.6f} oracle {s_o:.6f}")
    assert s[0] == s[1]
    assert abs(s[0] - s_o) <= SCORE_TOL


def test_mfma_channel_split_matches_whole_tiles():
    # a pair alone (379 tiles: one workgroup per tile and channel) against the
    # same pair in a batch of six (2,274 tiles: one workgroup per tile)
    clip = synth.video_clip(1920, 1080, 8, 6, 12)
    pairs = [clip.frame(f) for f in range(6)]
    clip.close()
    dp = api.DevicePairs(pairs, api.PF_YUV420_U8, 8, api.MATRIX_BT709)
    say("case: mfma map, 6 pairs 1920x1080 YUV420P8, each alone vs the batch of six")
    try:
        one = (api.Pair * 1)()
        alone = []
        with api.SSIMU2Batch(1, 1920, 1080) as b:
            for i in range(6):
                one[0] = dp.descs[i]
                alone.append(b.compute(one, 1, features=True))
        with api.SSIMU2Batch(6, 1920, 1080) as b:
            s, f = b.compute(dp.descs, features=True)
        for i in range(6):
            assert s[i] == alone[i][0][0], (i, s[i], alone[i][0][0])
            assert np.array_equal(f[i], alone[i][1][0]), i
        s_o, _ = ol.score_yuv420(pairs[0][0], pairs[0][1], 8, ol.SO_MAT_BT709)
        say(f"  scores {np.round(s, 4)}; pair 0 oracle {s_o:.6f}")
        assert abs(s[0] - s_o) <= SCORE_TOL
    finally:
        dp.free()
