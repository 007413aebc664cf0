"""The FFVship-style command line (vship_amd/ffvship) on YUV4MPEG2 files: every
score it writes against the FP64 oracle's score of the same frame pair
(|dscore| <= 1e-3; the JSON keeps six significant digits), across batch
boundaries and frame selections, in the reference's output formats."""
import json
import os
import subprocess

import numpy as np
import pytest

import vship_amd
from vship_amd import api
from tests import oracle_lib as ol
from tests.gpu_helpers import say
from tests.test_ffvship_cli import CLI, needs_cli, write_y4m
from tests.test_gpu_yuv_layouts import yuv_pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu_ready():
    vship_amd.gpu_full_check(0)
    yield


def clip(n, w, h, layout, bits, seed):
    pairs = [yuv_pair(w, h, layout, bits, seed + i) for i in range(n)]
    return [p[0] for p in pairs], [p[1] for p in pairs]


def run(*args):
    say("case: ffvship " + " ".join(os.path.basename(a) if a.endswith(".y4m") or a.endswith(".json") else a
                                    for a in args))
    r = subprocess.run([CLI, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout


def close(g, e):
    # 1e-3 parity plus the JSON's six significant digits
    return abs(g - e) <= 1e-3 + 5e-6 * abs(e)


def oracle_scores(ref, dist, bits, sw, sh, loc, frames):
    # the command's matrix: BT.709 above 650 rows, else BT.601 (all frames here are smaller)
    return [ol.score_yuv(ref[i], dist[i], bits, ol.SO_MAT_BT601, sw, sh, loc)[0] for i in frames]


@needs_cli
def test_420_clip_batches_selection_and_outputs(tmp_path):
    w, h, n = 160, 96, 11
    ref, dist = clip(n, w, h, "420", 8, 900)
    a, b, js = str(tmp_path / "ref.y4m"), str(tmp_path / "dist.y4m"), str(tmp_path / "s.json")
    write_y4m(a, ref, w, h, "C420mpeg2")
    write_y4m(b, dist, w, h, "C420mpeg2")
    expect = oracle_scores(ref, dist, 8, 1, 1, ol.SO_LOC_LEFT, range(n))
    out = run("--source", a, "--encoded", b, "-g", "1", "--json", js)  # batches of 8: crosses a boundary
    got = [x[0] for x in json.load(open(js))]
    say(f"  max |d| {max(abs(g - e) for g, e in zip(got, expect)):.1e}")
    assert len(got) == n and all(abs(g - e) <= 1e-3 for g, e in zip(got, expect))
    assert out.startswith(f"SSIMU2 Result between {a} and {b}\nComputed {n} frames at ")
    assert "-----------SSIMULACRA2-----------" in out and "95th percentile : " in out
    # the reference's median is the element n/2 of the sorted scores
    assert abs(float(out.split("Median : ")[1].split()[0]) - sorted(got)[n // 2]) < 1e-4

    out = run("-s", a, "-e", b, "--start", "1", "--every", "3", "--json", js, "--live-score-output")
    lines = out.strip().split("\n")
    assert lines[0] == "4"
    live = {int(l.split()[0]): float(l.split()[1]) for l in lines[1:]}
    assert sorted(live) == [0, 1, 2, 3]
    sel = [1, 4, 7, 10]
    got = [x[0] for x in json.load(open(js))]
    assert all(abs(got[k] - expect[f]) <= 1e-3 and abs(live[k] - expect[f]) <= 1e-3 for k, f in enumerate(sel))

    out = run("-s", a, "-e", b, "--encoded-offset", "1", "--json", js, "-t", "3")
    got = [x[0] for x in json.load(open(js))]
    cross = [ol.score_yuv(ref[i], dist[i + 1], 8, ol.SO_MAT_BT601, 1, 1, ol.SO_LOC_LEFT)[0] for i in range(n - 1)]
    say(f"  offset 1: json {got[:3]} oracle {cross[:3]}")
    assert len(got) == n - 1 and all(close(g, e) for g, e in zip(got, cross)), list(zip(got, cross))


@needs_cli
@pytest.mark.parametrize("layout,bits,tag,loc", [("422", 10, "C422p10", ol.SO_LOC_LEFT),
                                                 ("444", 8, "C444", ol.SO_LOC_LEFT),
                                                 ("420", 8, "C420jpeg", ol.SO_LOC_CENTER)])
def test_other_layouts(tmp_path, layout, bits, tag, loc):
    w, h, n = 120, 72, 3
    sw, sh = ol.LAYOUTS[layout]
    ref, dist = clip(n, w, h, layout, bits, 950)
    a, b, js = str(tmp_path / "ref.y4m"), str(tmp_path / "dist.y4m"), str(tmp_path / "s.json")
    write_y4m(a, ref, w, h, tag)
    write_y4m(b, dist, w, h, tag)
    run("-s", a, "-e", b, "--json", js)
    got = [x[0] for x in json.load(open(js))]
    expect = oracle_scores(ref, dist, bits, sw, sh, loc, range(n))
    say(f"  {tag}: max |d| {max(abs(g - e) for g, e in zip(got, expect)):.1e}")
    assert all(abs(g - e) <= 1e-3 for g, e in zip(got, expect))


@needs_cli
def test_butteraugli_metric(tmp_path):
    """-m Butteraugli: the three norms of every frame pair against the FP64
    oracle (YUV to R'G'B' as the SSIMULACRA2 front end, then the metric), with
    --qnorm and --intensity-target passed through; JSON [[q, 3, inf], ...],
    live lines "index q 3 inf" and one statistics block per norm."""
    from tests import butter_lib as bl

    w, h, n = 160, 96, 5
    ref, dist = clip(n, w, h, "420", 8, 950)
    a, b, js = str(tmp_path / "ref.y4m"), str(tmp_path / "dist.y4m"), str(tmp_path / "b.json")
    write_y4m(a, ref, w, h, "C420mpeg2")
    write_y4m(b, dist, w, h, "C420mpeg2")
    want = []
    for i in range(n):
        rgb = [ol.yuv_to_rgb_any(im, 8, ol.SO_MAT_BT601, 1, 1, ol.SO_LOC_LEFT).astype(np.float32)
               for im in (ref[i], dist[i])]
        want.append(bl.butteraugli(list(rgb[0]), list(rgb[1]), 3, 150.0))
    want = np.array(want)
    out = run("-s", a, "-e", b, "-m", "Butteraugli", "--qnorm", "3", "--intensity-target", "150", "-g", "1",
              "--json", js)
    got = np.array(json.load(open(js)))
    assert got.shape == (n, 3)
    # 1e-4 relative parity plus the JSON's six significant digits
    assert np.all(np.abs(got - want) <= 1e-4 * want + 5e-6 * np.abs(want)), (got, want)
    lines = out.splitlines()
    assert lines[0] == f"Butteraugli Result between {a} and {b}"
    # src/FFVship.cpp:496-498: "<q>-Norm", "3-Norm", "INF-Norm" blocks, in that order
    titles = [l.strip("-") for l in lines if l.startswith("-")]
    assert titles == ["3-Norm", "3-Norm", "INF-Norm"]
    live = run("-s", a, "-e", b, "-m", "butter", "--live-score-output").splitlines()
    assert live[0] == str(n)
    for i, l in enumerate(live[1:]):
        idx, q, n3, ninf = l.split()
        assert int(idx) == i
        w2 = bl.butteraugli(*[list(ol.yuv_to_rgb_any(im, 8, ol.SO_MAT_BT601, 1, 1, ol.SO_LOC_LEFT)
                                   .astype(np.float32)) for im in (ref[i], dist[i])])
        assert np.allclose([float(q), float(n3), float(ninf)], w2, rtol=2e-4)



@needs_cli
def test_butteraugli_420p10_clip(tmp_path):
    """-m Butteraugli on a C420p10 clip (left-located chroma in 16-bit
    containers: the staged 16-bit input kernel) against the FP64 oracle."""
    from tests import butter_lib as bl

    w, h, n = 144, 80, 3
    ref, dist = clip(n, w, h, "420", 10, 970)
    a, b, js = str(tmp_path / "ref.y4m"), str(tmp_path / "dist.y4m"), str(tmp_path / "b.json")
    write_y4m(a, ref, w, h, "C420p10")
    write_y4m(b, dist, w, h, "C420p10")
    want = np.array([bl.butteraugli(*[list(ol.yuv_to_rgb_any(im, 10, ol.SO_MAT_BT601, 1, 1, ol.SO_LOC_LEFT)
                                          .astype(np.float32)) for im in (ref[i], dist[i])])
                     for i in range(n)])
    run("-s", a, "-e", b, "-m", "Butteraugli", "--json", js)
    got = np.array(json.load(open(js)))
    say(f"  C420p10 butteraugli max rel {np.max(np.abs(got - want) / want):.1e}")
    assert got.shape == (n, 3)
    assert np.all(np.abs(got - want) <= 1e-4 * want + 5e-6 * np.abs(want)), (got, want)
