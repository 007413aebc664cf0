"""The benchmark workloads at their real sizes on the GPU, against the FP64
oracle:

* 3840x2160 YUV420P10 BT.2020 limited pairs with the XYB scales 0-1 stored in
  FP16 and with every plane in FP32; the per-frame |dscore| is printed and
  bound by 1e-3;
* the seeded 64-pair uneven batch (640x360 .. 3840x2160 YUV420P8) under the
  autotuned launch shape, eager and graph-captured: every pair bit-exact with
  the same pair scored alone under that shape, sampled pairs (one of them 4K)
  within 1e-3 of the oracle;
* the continuous-batching server with 1920x1080 8-bit and 3840x2160 10-bit
  clips among smaller ones: every clip equal to its frames scored alone, and
  sampled frames within 1e-3 of the oracle.
"""
import time

import numpy as np
import pytest

import vship_amd
from vship_amd import api, synth
from tests import oracle_lib as ol
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu

SCORE_TOL = 1e-3


@pytest.fixture(scope="module", autouse=True)
def gpu_ready():
    vship_amd.gpu_full_check(0)
    ol.set_threads(16)
    yield


def fmt_of(bits):
    if bits == 8:
        return api.PF_YUV420_U8, api.MATRIX_BT709, ol.SO_MAT_BT709
    return api.PF_YUV420_U16, api.MATRIX_BT2020NCL, ol.SO_MAT_BT2020NCL


def clip_pairs(w, h, bits, n, seed):
    c = synth.video_clip(w, h, bits, n, seed)
    pairs = [c.frame(f) for f in range(n)]
    c.close()
    return pairs


def oracle_score(pair, bits):
    return ol.score_yuv420(pair[0], pair[1], bits, fmt_of(bits)[2], "f64")[0]


def test_4k_10bit_fp16_and_fp32_planes():
    pairs = clip_pairs(3840, 2160, 10, 2, seed=2)
    fmt, matrix, _ = fmt_of(10)
    say("case: 2 pairs 3840x2160 YUV420P10 BT.2020, FP16 planes (scales 0-1) and FP32 planes")
    dp = api.DevicePairs(pairs, fmt, 10, matrix)
    try:
        with api.SSIMU2Batch(2, 3840, 2160, fp16=True) as b16:
            s16 = b16.compute(dp.descs)
        with api.SSIMU2Batch(2, 3840, 2160) as b32:
            s32 = b32.compute(dp.descs)
    finally:
        dp.free()
    want = np.array([oracle_score(p, 10) for p in pairs])
    for k in range(2):
        say(f"  frame {k}: oracle {want[k]:.6f} fp16 planes {s16[k]:.6f} (|d| {abs(s16[k] - want[k]):.2e}) "
            f"fp32 planes {s32[k]:.6f} (|d| {abs(s32[k] - want[k]):.2e})")
    assert np.all(np.abs(s16 - want) <= SCORE_TOL)
    assert np.all(np.abs(s32 - want) <= SCORE_TOL)


@pytest.fixture(scope="module")
def mixed64():
    res = synth.batch_resolutions(64, seed=3)
    clips, pairs = {}, []
    for i, (w, h) in enumerate(res):
        if (w, h) not in clips:
            clips[(w, h)] = synth.video_clip(w, h, 8, 64, seed=w + h)
        pairs.append(clips[(w, h)].frame(i))
    for c in clips.values():
        c.close()
    dp = api.DevicePairs(pairs, api.PF_YUV420_U8, 8, api.MATRIX_BT709)
    yield res, pairs, dp
    dp.free()


def test_mixed_64_tuned_and_graph(mixed64):
    res, pairs, dp = mixed64
    say(f"case: 64-pair uneven batch, {len(set(res))} resolutions up to 3840x2160, autotuned shape")
    t = api.autotune(dp.descs, 64, fp16=False, repeats=1)
    th, grid = t["map_tile_height"], t["persistent_grid"]
    say(f"  tuned: tile height {th} grid {grid} graph {t['use_graph']} {t['best_ms']:.3f} ms")
    with api.SSIMU2Batch(64, 3840, 2160, map_tile_height=th, persistent_grid=grid) as b:
        eager = b.compute(dp.descs)
    with api.SSIMU2Batch(64, 3840, 2160, map_tile_height=th, persistent_grid=grid, use_graph=True) as g:
        for rep in range(3):
            got = g.compute(dp.descs)
            assert np.array_equal(got, eager), rep
    # persistent grids (per-XCD dynamic tile queue), eager and captured: the
    # same scores bit for bit
    for pg in (8, 512, 2048):
        with api.SSIMU2Batch(64, 3840, 2160, map_tile_height=th, persistent_grid=pg) as b:
            assert np.array_equal(b.compute(dp.descs), eager), pg
        with api.SSIMU2Batch(64, 3840, 2160, map_tile_height=th, persistent_grid=pg, use_graph=True) as g:
            for rep in range(2):
                assert np.array_equal(g.compute(dp.descs), eager), (pg, rep)
    # each pair alone under the same shape
    one = (api.Pair * 1)()
    with api.SSIMU2Batch(1, 3840, 2160, map_tile_height=th, persistent_grid=grid) as b1:
        for i in range(64):
            one[0] = dp.descs[i]
            s = b1.compute(one, 1)[0]
            assert s == eager[i], (i, res[i], s, eager[i])
    # sampled pairs against the oracle: the first 4K pair and two others
    k4 = res.index((3840, 2160))
    sample = sorted({k4, 0, 63})
    for i in sample:
        want = oracle_score(pairs[i], 8)
        say(f"  pair {i} {res[i][0]}x{res[i][1]}: gpu {eager[i]:.6f} oracle {want:.6f} |d| {abs(eager[i] - want):.2e}")
        assert abs(eager[i] - want) <= SCORE_TOL


SERVED = [  # (w, h, bits, frames, seed)
    (1920, 1080, 8, 6, 31),
    (3840, 2160, 10, 4, 32),
    (640, 360, 10, 9, 33),
    (1280, 720, 8, 5, 34),
    (2560, 1440, 10, 3, 35),
]


def test_served_large_clips_equal_alone_and_oracle():
    host, dps = [], []
    for (w, h, bits, n, seed) in SERVED:
        pp = clip_pairs(w, h, bits, n, seed)
        fmt, matrix, _ = fmt_of(bits)
        host.append(pp)
        dps.append(api.DevicePairs(pp, fmt, bits, matrix))
    try:
        alone = []
        one = (api.Pair * 1)()
        with api.SSIMU2Batch(1, 3840, 2160) as b:
            for dp in dps:
                s = []
                for i in range(dp.n):
                    one[0] = dp.descs[i]
                    s.append(b.compute(one, 1)[0])
                alone.append(np.array(s))
        # the budget holds the 4K clip's pairs but forces the others to wait
        budget = 6 * vship_amd.pair_footprint(3840, 2160)
        say(f"case: server with 1080p 8-bit and 4K 10-bit clips, budget {budget / 1e9:.1f} GB, 8 pairs per step")
        with api.SSIMU2Server(budget, 8, 3840, 2160, max_inflight_clips=3) as srv:
            ids = [srv.submit(dp.descs) for dp in dps[:2]]
            time.sleep(0.005)
            ids += [srv.submit(dp.descs) for dp in dps[2:]]
            res = [srv.wait(i) for i in ids]
        for k, (scores, lat) in enumerate(res):
            say(f"  clip {k} {SERVED[k][:3]} latency {lat:.2f} ms")
            assert np.array_equal(scores, alone[k]), (k, scores, alone[k])
        # sampled frames against the oracle: first and last of the two large clips
        for k in (0, 1):
            bits = SERVED[k][2]
            for i in (0, SERVED[k][3] - 1):
                want = oracle_score(host[k][i], bits)
                got = res[k][0][i]
                say(f"  clip {k} frame {i}: served {got:.6f} oracle {want:.6f} |d| {abs(got - want):.2e}")
                assert abs(got - want) <= SCORE_TOL
    finally:
        for dp in dps:
            dp.free()
