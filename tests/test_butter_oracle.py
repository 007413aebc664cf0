"""Pins the Butteraugli oracle (oracle/butter_oracle.c).

* Against the reference's own HIP code. tests/golden/reference_butter.json
  holds the norms and a sample of the distortion map computed from src/butter
  unchanged, on the GPU. oracle/ref_butter_shim.hip ran that code, and
  tests/golden/make_golden.py ("reference-butter") wrote the file. This covers
  every constant and table of the metric: opsin, frequency bands, the 16+16
  Malta lines, L2 weights, masks and the scale blend.
* Against closed forms that do not restate the oracle's code:
  * the renormalised blur equals scipy's Gaussian filter of the image divided
    by the filter of the image's support;
  * identical images give a zero map;
  * the norms follow from the map by their definitions.
"""
import json
import os

import numpy as np
import pytest
from scipy import ndimage

from tests import butter_lib as bl
from tests.golden.make_golden import BUTTER_CASES, butter_inputs

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_butter.json")
SIGMAS = [1.2, 1.56416327805, 2.7, 3.22489901262, 7.15593339443]


def _golden():
    with open(GOLDEN) as fh:
        return {c["name"]: c for c in json.load(fh)["cases"]}


@pytest.mark.parametrize("si", range(5))
def test_blur_is_renormalised_gaussian(si):
    # src/butter/gaussianblur.hpp: window max(floor(3 sigma), 8); each output
    # divided by the sum of its in-image taps
    rng = np.random.default_rng(si)
    img = rng.random((29, 37))
    s = SIGMAS[si]
    radius = max(int(3 * s), 8)
    num = ndimage.gaussian_filter(img, s, mode="constant", cval=0.0, radius=radius)
    den = ndimage.gaussian_filter(np.ones_like(img), s, mode="constant", cval=0.0, radius=radius)
    np.testing.assert_allclose(bl.blur(img, si), num / den, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("si", [0, 4])
def test_blur_keeps_constants(si):
    img = np.full((17, 23), 0.37)
    np.testing.assert_allclose(bl.blur(img, si), img, rtol=1e-13)


def test_identical_images_give_zero():
    ref, _ = butter_inputs(64, 48, "f32", 11)
    norms, dm = bl.butteraugli(ref, [p.copy() for p in ref], 2, 203.0, diffmap=True)
    assert np.all(dm == 0.0)
    assert np.all(norms == 0.0)


@pytest.mark.parametrize("q", [1, 2, 3, 5])
def test_norms_follow_from_map(q):
    # src/butter/diffnorms.hpp:78-120: (mean |m|^q)^(1/q), (mean |m|^3)^(1/3), max |m|
    ref, dist = butter_inputs(40, 33, "f32", 5)
    norms, dm = bl.butteraugli(ref, dist, q, 203.0, diffmap=True)
    a = np.abs(dm)
    np.testing.assert_allclose(norms, [np.mean(a ** q) ** (1.0 / q), np.mean(a ** 3) ** (1.0 / 3), a.max()],
                               rtol=1e-12)


def test_malta_zero_for_equal_planes():
    rng = np.random.default_rng(3)
    p = rng.standard_normal((21, 19))
    for hf in (0, 1):
        assert np.all(bl.malta(p, p.copy(), hf, 2.0, 3.0, 5.0) == 0.0)


def test_malta_is_sum_of_squares():
    # every term is a squared line sum, so it is never negative. An impulse
    # difference d reaches each pixel through the lines that hold it: the map
    # is (number of such lines) * d^2, largest (16 d^2) at the impulse, which
    # all 16 lines pass through
    rng = np.random.default_rng(4)
    l1 = rng.standard_normal((25, 27))
    l0 = l1 + rng.standard_normal(l1.shape)
    for hf in (0, 1):
        assert np.all(bl.malta(l0, l1, hf, 1.5, 2.5, 7.0) >= 0.0)
        a = np.zeros((15, 15))
        a[7, 7] = 1e-3
        m = bl.malta(a, np.zeros_like(a), hf, 1.0, 1.0, 1e6)
        counts = m / (m[7, 7] / 16.0)
        np.testing.assert_allclose(counts, np.round(counts), atol=1e-9)
        assert m.max() == m[7, 7]
        assert np.count_nonzero(np.round(counts)) > 16


@pytest.mark.parametrize("case", [c[0] for c in BUTTER_CASES])
def test_oracle_matches_reference_hip(case):
    """The oracle (FP64) against the reference's own HIP code (FP32) on the
    same inputs: norms within 1e-5 relative (observed <= 1e-6), map samples
    within 1e-4 relative + 1e-6 absolute (observed <= 1.2e-5), map sums
    within 1e-5 relative."""
    g = _golden()[case]
    name, w, h, fmt, seed, q = next(c for c in BUTTER_CASES if c[0] == case)
    ref, dist = butter_inputs(w, h, fmt, seed)
    norms, dm = bl.butteraugli(ref, dist, q, 203.0, diffmap=True)
    np.testing.assert_allclose(norms, g["norms"], rtol=1e-5)
    got = np.array([dm[y, x] for x, y, _ in g["map_samples"]])
    want = np.array([v for _, _, v in g["map_samples"]])
    np.testing.assert_allclose(got, want, rtol=1e-4, atol=1e-6)
    assert dm.sum() == pytest.approx(g["map_sum"], rel=1e-5)


def _guard_refuses(n):
    """A width or height the reference's tiled blur reads out of bounds for:
    n or ceil(n / 2) leaves a remainder of 1..5 modulo 32
    (src/butter/gaussianblur.hpp:127-130,154-157)."""
    return any(1 <= m % 32 <= 5 for m in (n, (n + 1) // 2))


def test_reference_size_guard_python_side():
    """The test-side guard refuses exactly the sizes the reference's tiled
    blur reads out of bounds for, at full and half scale."""
    for w in range(1, 300):
        for h in (7, 16, 33, 64, 70, 97, 130):
            assert bl.ref_size_ok(w, h) == (not _guard_refuses(w) and not _guard_refuses(h)), (w, h)


@pytest.mark.skipif(not bl.ref_available(), reason="reference shim not built (reference sources absent)")
def test_reference_size_guard_in_shim():
    """vref_butter_size_ok, the guard inside the reference shim, refuses every
    such size, and vref_butter returns its error for one before touching the
    device (called here without a GPU)."""
    R = bl.ref_lib()
    for w in range(1, 300):
        for h in (7, 16, 33, 64, 70, 97, 130):
            assert bool(R.vref_butter_size_ok(w, h)) == (not _guard_refuses(w) and not _guard_refuses(h)), (w, h)
    refused = [n for n in range(1, 200) if _guard_refuses(n)]
    assert 33 in refused and 97 in refused and 65 in refused  # 65 -> 33 at half scale
    planes = [np.zeros((64, 97), np.float32) for _ in range(3)]
    with pytest.raises(ValueError):
        bl.ref_butteraugli(planes, planes)
    import ctypes
    ptrs = bl.P3(*[p.ctypes.data for p in planes])
    F = ctypes.POINTER(ctypes.c_float)
    norms = np.zeros(3, np.float32)
    rc = R.vref_butter(1, ptrs, ptrs, 97 * 4, 97 * 4, 97, 64, 2, 203.0, norms.ctypes.data_as(F),
                       ctypes.cast(None, F))
    assert rc == 3
