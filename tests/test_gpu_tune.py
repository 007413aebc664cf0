"""Autotuner of the batch launch shape (Vship_SSIMU2BatchTune): every candidate
shape runs, and the tuning table round-trips through a file into
Vship_SSIMU2BatchInit.

The map tile height sets how many rows one thread sums in FP32 before the
FP64 tile sum, so shapes of different tile height agree to rounding (1e-4 on
the score here); the persistent grid size and graph capture change nothing
(bit-exact)."""
import numpy as np
import pytest

import vship_amd
from vship_amd import api, synth
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def uneven():
    vship_amd.gpu_full_check(0)
    res = [(320, 180), (640, 360), (200, 120), (640, 360), (96, 64), (480, 270)]
    pairs = []
    for i, (w, h) in enumerate(res):
        c = synth.video_clip(w, h, 8, 2, seed=10 + i)
        pairs.append(c.frame(i % 2))
        c.close()
    dp = api.DevicePairs(pairs, api.PF_YUV420_U8, 8, api.MATRIX_BT709)
    with api.SSIMU2Batch(len(pairs), 640, 360, map_tile_height=48) as b:
        base = b.compute(dp.descs)
    yield dp, base
    dp.free()


def test_autotune_candidates_and_best(uneven):
    dp, base = uneven
    say("case: autotune over 6 uneven pairs up to 640x360")
    r = api.autotune(dp.descs, fp16=False, repeats=1)
    say(f"best th {r['map_tile_height']} grid {r['persistent_grid']} graph {r['use_graph']} {r['best_ms']:.3f} ms")
    assert len(r["candidates"]) >= 3
    assert r["map_tile_height"] in (1, 48, 64, 112)  # 1: matrix-core map kernel
    assert all(c["ms"] > 0 for c in r["candidates"])
    assert r["best_ms"] == pytest.approx(min(c["ms"] for c in r["candidates"]), rel=1e-6)
    with api.SSIMU2Batch(dp.n, 640, 360, map_tile_height=r["map_tile_height"],
                         persistent_grid=r["persistent_grid"], use_graph=bool(r["use_graph"])) as b:
        first = b.compute(dp.descs)
        assert np.max(np.abs(first - base)) <= 1e-4
        assert np.array_equal(b.compute(dp.descs), first)


@pytest.mark.parametrize("th,grid", [(64, 0), (112, 0), (48, 512), (64, 1024), (1, 0), (1, 512)])
def test_every_shape_scores_the_same(uneven, th, grid):
    dp, base = uneven
    say(f"case: shape th={th} grid={grid}, eager and graph")
    with api.SSIMU2Batch(dp.n, 640, 360, map_tile_height=th) as b:
        plain = b.compute(dp.descs)
    say(f"max |score(th={th}) - score(th=48)| = {np.max(np.abs(plain - base)):.2e}")
    assert np.max(np.abs(plain - base)) <= 1e-4
    with api.SSIMU2Batch(dp.n, 640, 360, map_tile_height=th, persistent_grid=grid) as b:
        assert np.array_equal(b.compute(dp.descs), plain)
    with api.SSIMU2Batch(dp.n, 640, 360, map_tile_height=th, persistent_grid=grid, use_graph=True) as b:
        for _ in range(3):
            assert np.array_equal(b.compute(dp.descs), plain)


def test_tune_table_round_trip(uneven, tmp_path):
    dp, base = uneven
    path = tmp_path / "tune.txt"
    api.autotune(dp.descs, repeats=1)
    api.tune_save(path)
    text = path.read_text().splitlines()
    say(f"tune table: {text}")
    rows = [l.split() for l in text if not l.startswith("#")]
    assert rows and all(len(r) == 6 for r in rows)
    api.tune_load(path)
    # a handle with the shape left at 0 picks up the tuned entry of its size class
    want = [int(r[2]) for r in rows if r[0] == "0" and r[1] == "0"]
    with api.SSIMU2Batch(dp.n, 640, 360) as b:
        s = b.compute(dp.descs)
        th = b.stats()["map_tile_height"]
    say(f"handle with tuned shape: th {th}, table {want}")
    assert want and th == want[0]
    assert np.max(np.abs(s - base)) <= 1e-4
    # put the default shape back for the size classes later tests use
    path.write_text("".join(f"{c} {f} 1 0 0 0\n" for c in range(5) for f in range(2)))
    api.tune_load(path)
    with api.SSIMU2Batch(dp.n, 640, 360) as b, api.SSIMU2Batch(dp.n, 640, 360, map_tile_height=1) as m:
        assert np.array_equal(b.compute(dp.descs), m.compute(dp.descs))
