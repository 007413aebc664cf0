"""Every planar YUV layout of the fused front end (4:2:0, 4:2:2, 4:4:4, 4:4:0,
4:1:1, 4:1:0, in 8-bit and 16-bit containers, with the four chroma
locations) through the batched C ABI, against the FP64 oracle's statement of
the same front end (|dscore| <= 1e-3, features as in the parity tests), and
a batch that mixes layouts against its pairs scored alone (bit-exact)."""
import numpy as np
import pytest

import vship_amd
from vship_amd import api
from tests import oracle_lib as ol
from tests.gpu_helpers import gpu_batch, rgb_pair, say

pytestmark = pytest.mark.gpu

SCORE_TOL = 1e-3
LOC_NAMES = {api.CHROMA_LEFT: "left", api.CHROMA_CENTER: "center", api.CHROMA_TOPLEFT: "topleft",
             api.CHROMA_TOP: "top"}


def feat_close(g, o):
    return np.all(np.abs(g - o) <= 2e-5 + 1e-3 * np.abs(o))


@pytest.fixture(scope="module", autouse=True)
def gpu_ready():
    vship_amd.gpu_full_check(0)
    yield


def yuv_pair(w, h, layout, bits, seed):
    """Smooth limited-range YUV frame of the layout and a noisy copy."""
    _, _, sw, sh = api.YUV_LAYOUTS[layout]
    rng = np.random.default_rng(seed)
    f, _ = rgb_pair(w, h, api.PF_RGB_F32, seed)
    s = 1 << (bits - 8)
    dt = np.uint8 if bits == 8 else np.uint16
    y = 16 * s + 219 * s * f[0].astype(np.float64)
    u = 16 * s + 224 * s * f[1].astype(np.float64)[::1 << sh, ::1 << sw]
    v = 16 * s + 224 * s * f[2].astype(np.float64)[::1 << sh, ::1 << sw]
    ref = [np.round(p).astype(dt) for p in (y, u, v)]
    dist = [np.clip(np.round(p + 3 * s * rng.standard_normal(p.shape)), 0, (1 << bits) - 1).astype(dt)
            for p in (y, u, v)]
    return ref, dist


CASES = [(layout, bits, loc) for layout in api.YUV_LAYOUTS for bits, loc in
         ((8, api.CHROMA_LEFT), (10, api.CHROMA_CENTER))] + \
        [(layout, 8, loc) for layout in ("420", "422", "410") for loc in (api.CHROMA_TOPLEFT, api.CHROMA_TOP)]


@pytest.mark.parametrize("layout,bits,loc", CASES)
def test_layout_against_oracle(layout, bits, loc):
    f8, f16, sw, sh = api.YUV_LAYOUTS[layout]
    fmt = f8 if bits == 8 else f16
    w, h = 203, 117
    ref, dist = yuv_pair(w, h, layout, bits, 700 + 3 * sw + sh)
    say(f"case: YUV{layout} {bits}-bit chroma {LOC_NAMES[loc]} {w}x{h}")
    matrix = api.MATRIX_BT709 if bits == 8 else api.MATRIX_BT2020NCL
    s_o, f_o = ol.score_yuv(ref, dist, bits, matrix, sw, sh, loc)
    (s_g,), (f_g,) = gpu_batch([(ref, dist)], fmt, bits=bits, matrix=matrix, chroma_location=loc)
    say(f"  oracle {s_o:.7f} gpu {s_g:.7f} ({s_g - s_o:+.1e})")
    assert abs(s_g - s_o) <= SCORE_TOL
    assert feat_close(f_g, f_o), np.max(np.abs(f_g - f_o))


def test_identical_frames_score_100_in_every_layout():
    for layout, (f8, _, _, _) in api.YUV_LAYOUTS.items():
        ref, _ = yuv_pair(97, 61, layout, 8, 5)
        say(f"case: YUV{layout} identical frames 97x61")
        (s,), (f,) = gpu_batch([(ref, ref)], f8, chroma_location=api.CHROMA_CENTER)
        assert s == 100.0 and np.all(f == 0)


def test_mixed_layout_batch_equals_pairs_alone():
    items = [("444", 8, api.CHROMA_LEFT, (160, 90)), ("420", 8, api.CHROMA_LEFT, (321, 181)),
             ("422", 10, api.CHROMA_TOP, (256, 144)), ("410", 8, api.CHROMA_CENTER, (99, 203)),
             ("420", 10, api.CHROMA_CENTER, (200, 120))]
    dps, alone = [], []
    for k, (layout, bits, loc, (w, h)) in enumerate(items):
        f8, f16, _, _ = api.YUV_LAYOUTS[layout]
        pair = yuv_pair(w, h, layout, bits, 40 + k)
        dp = api.DevicePairs([pair], f8 if bits == 8 else f16, bits, api.MATRIX_BT709, loc)
        dps.append(dp)
    say("case: one batch of 5 pairs in 5 layouts vs each pair alone")
    mw = max(it[3][0] for it in items)
    mh = max(it[3][1] for it in items)
    with api.SSIMU2Batch(len(items), mw, mh) as b:
        for dp in dps:
            alone.append(b.compute(dp.descs, 1)[0])
        mixed = (api.Pair * len(items))(*[dp.descs[0] for dp in dps])
        together = b.compute(mixed, len(items))
    for dp in dps:
        dp.free()
    assert np.array_equal(np.array(together), np.array(alone))


def test_bad_chroma_location_is_rejected():
    ref, dist = yuv_pair(64, 48, "422", 8, 1)
    dp = api.DevicePairs([(ref, dist)], api.PF_YUV422_U8, 8, api.MATRIX_BT709, 4)
    with api.SSIMU2Batch(1, 64, 48) as b:
        with pytest.raises(api.VshipError):
            b.compute(dp.descs, 1)
    dp.free()


def test_graph_captured_on_420_scores_later_mixed_layouts():
    # the captured graph holds both prep kernels; the generic one returns at
    # once for a batch without such pairs and does the work when one has them
    plain = api.DevicePairs([yuv_pair(160, 96, "420", 8, 60 + k) for k in range(2)], api.PF_YUV420_U8)
    other = [api.DevicePairs([yuv_pair(160, 96, lay, 8, 70 + k)], api.YUV_LAYOUTS[lay][0], 8, api.MATRIX_BT709,
                             api.CHROMA_CENTER) for k, lay in enumerate(("444", "411"))]
    mixed = (api.Pair * 3)(plain.descs[0], other[0].descs[0], other[1].descs[0])
    say("case: graph captured on a 4:2:0 batch, replayed on 4:2:0 + 4:4:4 + 4:1:1")
    with api.SSIMU2Batch(3, 160, 96) as b:
        eager = np.array(b.compute(mixed, 3))
        eager_plain = np.array(b.compute(plain.descs, 2))
    with api.SSIMU2Batch(3, 160, 96, use_graph=True) as g:
        assert np.array_equal(np.array(g.compute(plain.descs, 2)), eager_plain)
        for _ in range(3):
            assert np.array_equal(np.array(g.compute(mixed, 3)), eager)
            assert np.array_equal(np.array(g.compute(plain.descs, 2)), eager_plain)
    plain.free()
    for dp in other:
        dp.free()
