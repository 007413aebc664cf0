"""Pins for the CPU oracle (oracle/ssimu2_oracle.c).

Each stage is checked against something the mathematics fixes independently
of the oracle's own formula: closed forms, invariants, a different
formulation (direct 2-D normalised convolution, two-pass covariance,
edge-replicated block mean) or the reference's own known answer
(identical inputs with any row padding => exactly 100,
test/basic_metric_math_test.cpp:40-58 of the reference).
"""
import math
import os
import re

import numpy as np
import pytest

from tests import oracle_lib as ol

REF_SCORE_HPP = "/root/reference/src/ssimu2/score.hpp"


# ---------------------------------------------------------------- sizes
@pytest.mark.parametrize("w,h,T", [(256, 256, 87360), (1920, 1080, 2764200), (3840, 2160, 11056560),
                                   (640, 360, 307160), (1280, 720, 1228520), (2560, 1440, 4914000)])
def test_total_scale_size(w, h, T):
    assert ol.total_scale_size(w, h) == T


def test_total_scale_size_tiny():
    # always six scales, each ceil-halved, no early stop (src/ssimu2/main.hpp:34-42)
    assert ol.total_scale_size(1, 1) == 6
    assert ol.total_scale_size(3, 5) == 15 + 2 * 3 + 1 * 2 + 1 + 1 + 1


# ---------------------------------------------------------------- taps
def test_gaussian_taps():
    g = ol.gaussian_taps()
    assert np.allclose(g, g[::-1], rtol=0, atol=0)
    assert g[8] == pytest.approx(1 / math.sqrt(2 * math.pi * 2.25), rel=1e-15)
    for k in range(1, 9):
        assert g[8 + k] / g[8] == pytest.approx(math.exp(-k * k / 4.5), rel=1e-14)


# ---------------------------------------------------------------- linearise
def test_linearize_closed_form():
    assert ol.linearize(0.0) == 0.0
    assert ol.linearize(1.0) == 1.0
    assert ol.linearize(2.0) == pytest.approx(2 ** 2.4, rel=1e-15)
    for v in (0.1, 0.5, 0.73):
        assert ol.linearize(-v) == -ol.linearize(v)
        assert ol.linearize(v) == pytest.approx(math.exp(2.4 * math.log(v)), rel=1e-14)


# ---------------------------------------------------------------- XYB
def test_xyb_black():
    # cbrt(bias) == -abs_bias to ~1e-9, so black maps to the positive-XYB offsets
    x = ol.xyb([0.0, 0.0, 0.0])
    assert x == pytest.approx([0.42, 0.01, 0.55], abs=1e-8)


@pytest.mark.parametrize("g", [0.0, 0.01, 0.18, 0.5, 1.0, 1.7])
def test_xyb_grey_axis(g):
    # every opsin row sums to one, so grey has X = 0.42 and B = 0.55 (a transposed
    # matrix or a wrong row breaks this), and Y follows the cube root closed form
    x = ol.xyb([g, g, g])
    assert x[0] == pytest.approx(0.42, abs=1e-7)
    assert x[2] == pytest.approx(0.55, abs=1e-7)
    assert x[1] == pytest.approx((g + 0.0037930734) ** (1 / 3) - 0.1559542025327239 + 0.01, abs=1e-7)


def test_xyb_red_green_order():
    # X is 14 * (l - m) / 2: pure red raises l more than m (0.30 vs 0.23), pure
    # green lowers it (0.622 vs 0.692); B grows with blue.
    r, gr, b = ol.xyb([1, 0, 0]), ol.xyb([0, 1, 0]), ol.xyb([0, 0, 1])
    assert r[0] > 0.42 > gr[0]
    assert b[2] > 0.55 > r[2]


# ---------------------------------------------------------------- downsample
def _downsample_independent(a):
    # edge-replicate to even size, then 2x2 block mean: equal to the clamped index rule
    h, w = a.shape
    p = np.pad(a, ((0, h % 2), (0, w % 2)), mode="edge")
    return p.reshape(p.shape[0] // 2, 2, p.shape[1] // 2, 2).mean(axis=(1, 3))


@pytest.mark.parametrize("h,w", [(1, 1), (3, 5), (5, 3), (4, 6), (7, 9), (16, 1), (1, 17)])
def test_downsample_brute(h, w):
    rng = np.random.default_rng(h * 100 + w)
    a = rng.random((h, w))
    assert np.allclose(ol.downsample(a), _downsample_independent(a), rtol=0, atol=1e-15)


def test_downsample_ramp():
    a = np.tile(np.arange(9, dtype=float), (4, 1))
    d = ol.downsample(a)
    assert np.allclose(d[0, :4], 2 * np.arange(4) + 0.5)
    assert d[0, 4] == 8.0  # odd width: the last child is the clamped last column


# ---------------------------------------------------------------- blur
def _blur_direct_2d(a):
    """Direct 2-D normalised convolution over the in-bounds 17x17 window."""
    h, w = a.shape
    k = np.arange(-8, 9)
    g = np.exp(-k * k / 4.5)
    out = np.zeros_like(a)
    for y in range(h):
        for x in range(w):
            num = den = 0.0
            for j in range(17):
                yy = y + j - 8
                if not 0 <= yy < h:
                    continue
                for i in range(17):
                    xx = x + i - 8
                    if not 0 <= xx < w:
                        continue
                    wgt = g[j] * g[i]
                    num += wgt * a[yy, xx]
                    den += wgt
            out[y, x] = num / den
    return out


@pytest.mark.parametrize("h,w", [(1, 1), (2, 19), (13, 11), (20, 23)])
def test_blur_matches_direct_2d(h, w):
    rng = np.random.default_rng(7 + h * w)
    a = rng.random((h, w))
    assert np.allclose(ol.blur(a), _blur_direct_2d(a), rtol=0, atol=1e-13)


def test_blur_constant_everywhere():
    a = np.full((37, 29), 0.731)
    assert np.allclose(ol.blur(a), 0.731, rtol=0, atol=1e-14)


def test_blur_impulse_interior():
    a = np.zeros((41, 41))
    a[20, 20] = 1.0
    k = np.arange(-8, 9)
    g = np.exp(-k * k / 4.5)
    g /= g.sum()
    b = ol.blur(a)
    assert np.allclose(b[12:29, 12:29], np.outer(g, g), rtol=0, atol=1e-16)
    assert b[11, 20] == 0.0 and b[20, 29] == 0.0


# ---------------------------------------------------------------- maps
def _maps_two_pass(i1, i2):
    """SSIM/edge maps with centred (two-pass) Gaussian moments."""
    h, w = i1.shape
    k = np.arange(-8, 9)
    g = np.exp(-k * k / 4.5)
    d = np.zeros_like(i1)
    art = np.zeros_like(i1)
    det = np.zeros_like(i1)
    for y in range(h):
        for x in range(w):
            ys = slice(max(0, y - 8), min(h, y + 9))
            xs = slice(max(0, x - 8), min(w, x + 9))
            wy = g[ys.start - y + 8: ys.stop - y + 8]
            wx = g[xs.start - x + 8: xs.stop - x + 8]
            W = np.outer(wy, wx)
            W /= W.sum()
            a, b = i1[ys, xs], i2[ys, xs]
            mu1, mu2 = (W * a).sum(), (W * b).sum()
            s11 = (W * (a - mu1) ** 2).sum()
            s22 = (W * (b - mu2) ** 2).sum()
            s12 = (W * (a - mu1) * (b - mu2)).sum()
            ssim = (1 - (mu1 - mu2) ** 2) * (2 * s12 + 0.0009) / (s11 + s22 + 0.0009)
            d[y, x] = max(1 - ssim, 0)
            v = (abs(i2[y, x] - mu2) + 1) / (abs(i1[y, x] - mu1) + 1) - 1
            art[y, x], det[y, x] = max(v, 0), max(-v, 0)
    return d, art, det


def test_maps_match_two_pass_moments():
    rng = np.random.default_rng(3)
    i1 = 0.3 + 0.5 * rng.random((14, 17))
    i2 = np.clip(i1 + 0.2 * rng.standard_normal(i1.shape), 0, 1)
    got = ol.channel_maps(i1, i2)
    exp = _maps_two_pass(i1, i2)
    for a, b in zip(got, exp):
        assert np.allclose(a, b, rtol=0, atol=1e-12)
    assert got[0].max() > 0.01 and got[1].max() > 0.01 and got[2].max() > 0.01


def test_maps_identical_are_zero():
    rng = np.random.default_rng(4)
    i = rng.random((19, 21))
    for m in ol.channel_maps(i, i):
        assert np.all(m == 0.0)


def test_maps_constants_closed_form():
    # two constant images: mu = value, sigma = 0 -> d = (c1 - c2)^2, edge maps 0
    d, art, det = ol.channel_maps(np.full((9, 12), 0.6), np.full((9, 12), 0.45))
    assert np.allclose(d, 0.15 ** 2, rtol=0, atol=1e-12)
    assert np.allclose(art, 0, atol=1e-14) and np.allclose(det, 0, atol=1e-14)


def test_scale_features_norms():
    rng = np.random.default_rng(5)
    x1 = rng.random((3, 10, 12))
    x2 = rng.random((3, 10, 12))
    f = ol.scale_features(x1, x2)
    for c in range(3):
        maps = ol.channel_maps(x1[c], x2[c])
        for m in range(3):
            assert f[c, m, 0] == pytest.approx(maps[m].mean(), rel=1e-13)
            assert f[c, m, 1] == pytest.approx((maps[m] ** 4).mean() ** 0.25, rel=1e-13)


# ---------------------------------------------------------------- score
def _ref_weights():
    src = open(REF_SCORE_HPP).read()
    blk = src[src.index("const float weights[108]"):]
    blk = blk[blk.index("{") + 1: blk.index("};")]
    return [float(v.strip().rstrip("f")) for v in blk.split(",") if v.strip()]


def _poly_inverse_weight(score):
    # invert score = 100 - 10 * P(0.9562... * S)^0.6276... for S = w * t
    q = ((100.0 - score) / 10.0) ** (1 / 0.6276336467831387)
    lo, hi = 0.0, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        p = 6.248496625763138e-5 * mid ** 3 + 2.326765642916932 * mid - 0.020884521182843837 * mid ** 2
        lo, hi = (mid, hi) if p < q else (lo, mid)
    return lo / 0.9562382616834844


def test_final_score_zero_is_100():
    assert ol.final_score(np.zeros(108)) == 100.0


def test_final_score_monotone():
    f = np.zeros(108)
    prev = 100.0
    for t in (1e-4, 1e-3, 1e-2, 5e-2):
        f[48] = t  # Y channel, scale 2, 1-norm SSIM
        s = ol.final_score(f)
        assert s < prev
        prev = s


@pytest.mark.skipif(not os.path.exists(REF_SCORE_HPP), reason="reference sources not mounted")
def test_final_score_weights_match_reference_table():
    # recover each weight from the score of a unit feature vector and compare
    # with the reference's own table, index by index
    w_ref = _ref_weights()
    assert len(w_ref) == 108
    for i in range(108):
        f = np.zeros(108)
        f[i] = 1e-3
        s = ol.final_score(f)
        if w_ref[i] == 0.0:
            assert s == 100.0
        else:
            assert _poly_inverse_weight(s) / 1e-3 == pytest.approx(w_ref[i], rel=1e-6), i


# ---------------------------------------------------------------- end to end
def test_identical_padded_zero_images_score_exactly_100():
    # the reference's known answer: zero images, one with 0xFF row padding
    w, h, stride = 57, 41, 64 * 4
    a = [np.zeros(stride * h, np.uint8) for _ in range(3)]
    b = [np.zeros(stride * h, np.uint8) for _ in range(3)]
    for p in b:
        p.reshape(h, stride)[:, w * 4:] = 255
    s, f = ol.score_rgb_strided(ol.SO_RGB_F32, a, stride, b, stride, w, h)
    assert s == 100.0
    assert np.all(f == 0.0)


def test_identical_random_images_score_exactly_100():
    rng = np.random.default_rng(11)
    img = [rng.integers(0, 256, (33, 47), dtype=np.uint8) for _ in range(3)]
    s, _ = ol.score_rgb(ol.SO_RGB_U8, img, [p.copy() for p in img])
    assert s == 100.0


def test_constant_images_closed_form_score():
    # two flat grey-level images: every scale keeps the constants, so each
    # feature is (dX_c)^2 for SSIM (1- and 4-norm alike) and 0 for the edge maps
    w, h = 40, 24
    ga, gb = 0.6, 0.35
    a = np.full((3, h, w), ga)
    b = np.full((3, h, w), gb)
    s, f = ol.score_linear(a, b)
    xa, xb = ol.xyb([ga] * 3), ol.xyb([gb] * 3)
    exp = np.zeros(108)
    for c in range(3):
        for sc in range(6):
            exp[c * 36 + sc * 6 + 0] = (xa[c] - xb[c]) ** 2
            exp[c * 36 + sc * 6 + 3] = (xa[c] - xb[c]) ** 2
    assert np.allclose(f, exp, rtol=1e-9, atol=1e-15)
    assert s == pytest.approx(ol.final_score(exp), abs=1e-8)
    assert s < 100.0


def test_input_formats_agree():
    # the same picture in u8, u16, f16 and f32 gives the same score to the
    # precision of each format
    rng = np.random.default_rng(12)
    a8 = [rng.integers(0, 256, (24, 30), dtype=np.uint8) for _ in range(3)]
    b8 = [np.clip(p.astype(int) + rng.integers(-20, 21, p.shape), 0, 255).astype(np.uint8) for p in a8]
    s8, _ = ol.score_rgb(ol.SO_RGB_U8, a8, b8)
    s32, _ = ol.score_rgb(ol.SO_RGB_F32, [p.astype(np.float32) / 255 for p in a8],
                          [p.astype(np.float32) / 255 for p in b8])
    s16, _ = ol.score_rgb(ol.SO_RGB_U16, [p.astype(np.uint16) * 257 for p in a8],
                          [p.astype(np.uint16) * 257 for p in b8])
    sh, _ = ol.score_rgb(ol.SO_RGB_F16, [(p.astype(np.float32) / 255).astype(np.float16) for p in a8],
                         [(p.astype(np.float32) / 255).astype(np.float16) for p in b8])
    assert s32 == pytest.approx(s8, abs=1e-5)
    assert s16 == pytest.approx(s8, abs=1e-5)
    assert sh == pytest.approx(s8, abs=0.05)


def test_f32_twin_close_to_f64():
    rng = np.random.default_rng(13)
    a = [rng.integers(0, 256, (64, 72), dtype=np.uint8) for _ in range(3)]
    b = [np.clip(p.astype(int) + rng.integers(-6, 7, p.shape), 0, 255).astype(np.uint8) for p in a]
    s64, f64 = ol.score_rgb(ol.SO_RGB_U8, a, b, "f64")
    s32, f32 = ol.score_rgb(ol.SO_RGB_U8, a, b, "f32")
    assert s32 == pytest.approx(s64, abs=2e-2)
    assert 0 < s64 < 100


# ---------------------------------------------------------------- YUV front end
def _kr_kb(matrix):
    return {0: (0.2126, 0.0722), 1: (0.299, 0.114), 2: (0.2627, 0.0593)}[matrix]


@pytest.mark.parametrize("bits,matrix", [(8, 0), (10, 2), (8, 1), (12, 0)])
def test_yuv_constant_planes(bits, matrix):
    w, h = 13, 9
    s = 1 << (bits - 8)
    yv, uv, vv = 100 * s, 90 * s, 170 * s
    dt = np.uint8 if bits == 8 else np.uint16
    y = np.full((h, w), yv, dt)
    u = np.full(((h + 1) // 2, (w + 1) // 2), uv, dt)
    v = np.full(((h + 1) // 2, (w + 1) // 2), vv, dt)
    rgb = ol.yuv_to_rgb([y, u, v], bits, matrix)
    kr, kb = _kr_kb(matrix)
    Y = (yv - 16 * s) / (219 * s)
    cb = (uv - 128 * s) / (224 * s)
    cr = (vv - 128 * s) / (224 * s)
    R = Y + 2 * (1 - kr) * cr
    B = Y + 2 * (1 - kb) * cb
    G = (Y - kr * R - kb * B) / (1 - kr - kb)
    assert np.allclose(rgb[0], R, atol=1e-13)
    assert np.allclose(rgb[1], G, atol=1e-13)
    assert np.allclose(rgb[2], B, atol=1e-13)


def test_yuv_chroma_siting_linear_ramps():
    # Catmull-Rom reproduces linear functions, so away from the clamped edges a
    # ramp in chroma lands exactly at the sited position: left-sited
    # horizontally (luma x <-> chroma x/2), centre-sited vertically
    # (luma y <-> chroma (y - 0.5)/2).
    w, h = 32, 24
    cw, ch = w // 2, h // 2
    y = np.full((h, w), 128, np.uint8)
    jx = np.arange(cw)
    iy = np.arange(ch)
    u = (np.tile(100 + 8 * jx, (ch, 1))).astype(np.uint8)           # ramp along x
    v = (np.tile(100 + 8 * iy[:, None], (1, cw))).astype(np.uint8)  # ramp along y
    rgb = ol.yuv_to_rgb([y, u, v], 8, 0)
    kr, kb = _kr_kb(0)
    Y = (128 - 16) / 219
    # B = Y + 2(1-kb) cb  ->  cb from B
    cb = (rgb[2] - Y) / (2 * (1 - kb))
    cr = (rgb[0] - Y) / (2 * (1 - kr))
    xs = np.arange(4, w - 6)
    ys = np.arange(4, h - 6)
    exp_cb = (100 + 8 * (xs / 2.0) - 128) / 224
    exp_cr = (100 + 8 * ((ys - 0.5) / 2.0) - 128) / 224
    assert np.allclose(cb[5, xs], exp_cb, atol=1e-13)
    assert np.allclose(cr[ys, 7], exp_cr, atol=1e-13)


def _hermite_weights(t):
    # cubic Hermite basis with central-difference tangents, as weights on p[-1..2]
    h00 = 2 * t ** 3 - 3 * t ** 2 + 1
    h10 = t ** 3 - 2 * t ** 2 + t
    h01 = -2 * t ** 3 + 3 * t ** 2
    h11 = t ** 3 - t ** 2
    return np.array([-h10 / 2, h00 - h11 / 2, h01 + h10 / 2, h11 / 2])


def test_yuv_catmull_rom_random_interior():
    rng = np.random.default_rng(21)
    w, h = 20, 16
    cw, ch = w // 2, h // 2
    y = rng.integers(16, 236, (h, w), dtype=np.uint8)
    u = rng.integers(16, 241, (ch, cw), dtype=np.uint8)
    v = rng.integers(16, 241, (ch, cw), dtype=np.uint8)
    rgb = ol.yuv_to_rgb([y, u, v], 8, 0)
    kb = 0.0722
    Y = (y.astype(float) - 16) / 219
    cb = (rgb[2] - Y) / (2 * (1 - kb)) * 224 + 128
    wx = {0: np.array([0, 1.0, 0, 0]), 1: _hermite_weights(0.5)}
    for yy in range(h):
        t = 0.25 if yy % 2 else 0.75
        i0 = (yy - 1) // 2 if yy % 2 else yy // 2 - 1
        wy = _hermite_weights(t)
        for xx in range(w):
            k0 = xx // 2 if xx % 2 == 0 else (xx - 1) // 2
            acc = 0.0
            for a in range(4):
                ry = min(max(i0 - 1 + a, 0), ch - 1)
                for b in range(4):
                    rx = min(max(k0 - 1 + b, 0), cw - 1)
                    acc += wy[a] * wx[xx % 2][b] * u[ry, rx]
            assert cb[yy, xx] == pytest.approx(acc, abs=1e-9)


def test_yuv16_samples_above_depth_are_clamped():
    # 16-bit-container samples above the declared depth's largest code are
    # clamped to it (the front end's definition; it bounds the XYB range)
    rng = np.random.default_rng(5)
    w, h = 24, 16
    planes = [rng.integers(64, 941, (h, w)).astype(np.uint16),
              rng.integers(64, 961, (h // 2, w // 2)).astype(np.uint16),
              rng.integers(64, 961, (h // 2, w // 2)).astype(np.uint16)]
    wild = [p.copy() for p in planes]
    wild[0][3, 5] = 65535
    wild[1][2, 2] = 4000
    clamped = [p.copy() for p in wild]
    for p in clamped:
        np.minimum(p, 1023, out=p)
    a = ol.yuv_to_rgb(wild, 10, ol.SO_MAT_BT2020NCL)
    b = ol.yuv_to_rgb(clamped, 10, ol.SO_MAT_BT2020NCL)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, ol.yuv_to_rgb(planes, 10, ol.SO_MAT_BT2020NCL))


# ------------------------------------------------- YUV front end: every layout
def _ref_chroma_coords(f, centred, n_chroma):
    """Luma index -> chroma coordinate, enumerated from the writes of the
    reference's upsampling kernels (src/ffvship_utility/gpuColorToLinear/
    chromaUpsample.hpp): left/top x2 write dst[2x] <- t=0, dst[2x+1] <- t=.5
    of interval x (:58-68, :116-128); left x4 write dst[4x+k] <- t=k/4
    (:86-98); centre x2 write dst[2x+1] <- t=.25, dst[2x+2] <- t=.75 (:44-56,
    :100-114); centre x4 write dst[4x+2..4x+5] <- t=1/8, 3/8, 5/8, 7/8
    (:70-84). No subsampling: luma x is chroma x."""
    coords = {}
    for x in range(-1, n_chroma):
        if f == 1:
            coords[x] = float(x)
        elif not centred:
            for k in range(f):
                coords[f * x + k] = x + k / f
        elif f == 2:
            coords[2 * x + 1], coords[2 * x + 2] = x + 0.25, x + 0.75
        else:
            for k, t in zip(range(2, 6), (0.125, 0.375, 0.625, 0.875)):
                coords[4 * x + k] = x + t
    return coords


@pytest.mark.parametrize("layout", list(ol.LAYOUTS))
@pytest.mark.parametrize("loc", [ol.SO_LOC_LEFT, ol.SO_LOC_CENTER, ol.SO_LOC_TOPLEFT, ol.SO_LOC_TOP])
def test_yuv_any_layout_quadratic_chroma_lands_at_reference_positions(layout, loc):
    # Catmull-Rom reproduces quadratics (central differences are exact for
    # them), so away from the clamped edges a quadratic chroma field comes
    # back exactly at the chroma coordinate of each luma sample
    sw, sh = ol.LAYOUTS[layout]
    fx, fy = 1 << sw, 1 << sh
    w, h = 48 + 3, 40 + 1
    cw, ch = -(-w // fx), -(-h // fy)
    hc = loc in (ol.SO_LOC_CENTER, ol.SO_LOC_TOP)
    vc = loc in (ol.SO_LOC_LEFT, ol.SO_LOC_CENTER)
    q = lambda i, j: 300.0 + 7.0 * i - 5.0 * j + 0.25 * i * i + 0.5 * j * j
    jj, ii = np.meshgrid(np.arange(cw), np.arange(ch))
    u = (q(jj, ii) * 16).astype(np.uint16)  # integer codes below 2^16; read as bits = 16
    assert u.max() < 65535 and np.array_equal(u, q(jj, ii) * 16)
    cx = _ref_chroma_coords(fx, hc, cw)
    cy = _ref_chroma_coords(fy, vc, ch)
    checked = 0
    for y in range(h):
        for x in range(w):
            px, py = cx[x], cy[y]
            if not (1 <= px <= cw - 3 and 1 <= py <= ch - 3):
                continue  # a tap would be clamped
            got = ol.chroma_sample(u, w, h, 16, sw, sh, loc, x, y)
            assert got == pytest.approx(q(px, py) * 16, abs=1e-7), (x, y, px, py)
            checked += 1
    assert checked > 200


@pytest.mark.parametrize("layout", list(ol.LAYOUTS))
def test_yuv_any_layout_constant_planes(layout):
    sw, sh = ol.LAYOUTS[layout]
    w, h = 17, 11
    cw, ch = -(-w // (1 << sw)), -(-h // (1 << sh))
    for loc in range(4):
        y = np.full((h, w), 150, np.uint8)
        u = np.full((ch, cw), 60, np.uint8)
        v = np.full((ch, cw), 200, np.uint8)
        rgb = ol.yuv_to_rgb_any([y, u, v], 8, 0, sw, sh, loc)
        kr, kb = _kr_kb(0)
        Y, cb, cr = (150 - 16) / 219, (60 - 128) / 224, (200 - 128) / 224
        R, B = Y + 2 * (1 - kr) * cr, Y + 2 * (1 - kb) * cb
        assert np.allclose(rgb[0], R, atol=1e-13) and np.allclose(rgb[2], B, atol=1e-13)
        assert np.allclose(rgb[1], (Y - kr * R - kb * B) / (1 - kr - kb), atol=1e-13)


def test_yuv444_takes_chroma_as_is():
    rng = np.random.default_rng(8)
    w, h = 19, 7
    y, u, v = (rng.integers(16, 241, (h, w), dtype=np.uint8) for _ in range(3))
    for loc in range(4):
        rgb = ol.yuv_to_rgb_any([y, u, v], 8, 1, 0, 0, loc)
        kr, kb = _kr_kb(1)
        Y = (y - 16.0) / 219
        assert np.allclose(rgb[0], Y + 2 * (1 - kr) * (v - 128.0) / 224, atol=1e-13)
        assert np.allclose(rgb[2], Y + 2 * (1 - kb) * (u - 128.0) / 224, atol=1e-13)


def test_yuv_any_layout_420_left_is_the_420_front_end():
    rng = np.random.default_rng(31)
    for w, h in [(21, 13), (32, 18), (1, 1), (2, 3)]:
        cw, ch = (w + 1) // 2, (h + 1) // 2
        planes = [rng.integers(0, 1024, (h, w)).astype(np.uint16),
                  rng.integers(0, 1024, (ch, cw)).astype(np.uint16),
                  rng.integers(0, 1024, (ch, cw)).astype(np.uint16)]
        assert np.array_equal(ol.yuv_to_rgb_any(planes, 10, 2, 1, 1, ol.SO_LOC_LEFT), ol.yuv_to_rgb(planes, 10, 2))


def test_yuv411_centre_random_interior_brute_force():
    # 4:1:1 with centred chroma: fractions 1/8 .. 7/8 of the Hermite basis
    rng = np.random.default_rng(12)
    w, h = 40, 5
    cw = w // 4
    u = rng.integers(16, 241, (h, cw), dtype=np.uint8)
    for x in range(w):
        c = (x - 1.5) / 4
        k0 = int(np.floor(c))
        wts = _hermite_weights(c - k0)
        exp = sum(wts[b] * u[3, min(max(k0 - 1 + b, 0), cw - 1)] for b in range(4))
        assert ol.chroma_sample(u, w, h, 8, 2, 0, ol.SO_LOC_CENTER, x, 3) == pytest.approx(exp, abs=1e-9)
