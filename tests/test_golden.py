"""The oracle against the fixtures of tests/golden (CPU only).

* oracle_vectors.json (written by the oracle itself): a regression record
  of the config-0 score and features, the scale sizes, the taps and a YUV
  4:2:0 sample; recomputed here to 1e-12.
* reference_hip.json: the reference's own HIP code's score and features on
  seeded inputs (tests/golden/make_golden.py reference, run on a GPU). The
  reference computes in FP32 with FP32 host sums; its features must lie in
  the band around the FP64 oracle that FP32 rounding explains
  (oracle_lib.reference_feature_mismatch), its score within 1e-2.
"""
import json
import os

import numpy as np
import pytest

from tests import oracle_lib as ol
from tests.golden.make_golden import FMT_ORACLE, case_inputs, yuv_sample

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    p = os.path.join(HERE, name)
    if not os.path.exists(p):
        pytest.skip(f"{name} not present")
    with open(p) as fh:
        return json.load(fh)


def test_oracle_config0_vector():
    from vship_amd import synth

    g = load("oracle_vectors.json")["config0"]
    ref, dist = synth.config0()
    chk = [int(np.sum(p.astype(np.uint64) * (np.arange(p.size, dtype=np.uint64) % 251 + 1).reshape(p.shape)))
           for p in ref + dist]
    assert chk == g["input_sha"], "synthetic config-0 input changed"
    s, f = ol.score_rgb(ol.SO_RGB_U8, ref, dist, "f64")
    assert s == pytest.approx(g["score"], rel=1e-12, abs=1e-12)
    assert np.allclose(f, g["features"], rtol=1e-11, atol=1e-15)


def test_oracle_small_vectors():
    g = load("oracle_vectors.json")
    assert [list(d) for d in ol.scale_dims(1920, 1080)] == g["scale_dims_1920x1080"]
    assert np.array_equal(ol.gaussian_taps(), np.array(g["gaussian_taps"]))
    y = g["yuv420p10_bt2020_8x6"]
    planes = yuv_sample()
    assert [p.tolist() for p in planes] == y["planes"]
    rgb = ol.yuv_to_rgb(planes, 10, ol.SO_MAT_BT2020NCL)
    assert np.allclose(rgb, np.array(y["rgb"]), rtol=0, atol=1e-14)


def reference_cases():
    p = os.path.join(HERE, "reference_hip.json")
    if not os.path.exists(p):
        return []
    with open(p) as fh:
        return json.load(fh)["cases"]


@pytest.mark.parametrize("case", reference_cases(), ids=lambda c: c["name"])
def test_oracle_matches_reference_hip_output(case):
    ref, dist = case_inputs(case["w"], case["h"], case["format"], case["seed"])
    s_o, f_o = ol.score_rgb(FMT_ORACLE[case["format"]], ref, dist, "f64")
    bad = ol.reference_feature_mismatch(case["features"], f_o)
    assert not bad, [(i, case["features"][i], f_o[i]) for i in bad]
    assert abs(case["score"] - s_o) <= 1e-2


def test_reference_fixture_present():
    if not reference_cases():
        pytest.skip("reference_hip.json not generated yet")
    assert {c["name"] for c in reference_cases()} >= {"config0"}


def test_reference_band_rejects_planted_mistakes():
    # plausible oracle mistakes, planted in the oracle's own output on the CPU,
    # must leave the band around the reference's features
    cases = reference_cases()
    if not cases:
        pytest.skip("reference_hip.json not generated yet")
    c = cases[0]
    ref, dist = case_inputs(c["w"], c["h"], c["format"], c["seed"])
    _, f = ol.score_rgb(FMT_ORACLE[c["format"]], ref, dist, "f64")
    assert not ol.reference_feature_mismatch(c["features"], f)
    idx = np.arange(108)
    m, n, s = idx % 3, ((idx % 36) % 6) // 3, (idx % 36) // 6
    swapped_maps = f.copy()
    swapped_maps[m == 1], swapped_maps[m == 2] = f[m == 2], f[m == 1]      # artifact <-> detail
    swapped_scales = f.copy()
    swapped_scales[s == 0], swapped_scales[s == 1] = f[s == 1], f[s == 0]  # scale order
    wrong_root = f.copy()
    wrong_root[n == 1] = f[n == 1] ** 2                                    # sqrt instead of 4th root
    swapped_channels = np.concatenate([f[36:72], f[:36], f[72:]])          # X <-> Y
    for name, bad in [("maps", swapped_maps), ("scales", swapped_scales), ("root", wrong_root),
                      ("channels", swapped_channels)]:
        assert ol.reference_feature_mismatch(c["features"], bad), name
