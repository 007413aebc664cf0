"""Continuous-batching server: every clip's scores equal the same frames
scored alone (bit for bit), under a memory budget that forces admission to
wait, with clips arriving while others run."""
import time

import numpy as np
import pytest

import vship_amd
from vship_amd import api, synth
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu

CLIPS = [  # (w, h, bits, frames, seed)
    (160, 90, 8, 5, 1),
    (320, 180, 10, 3, 2),
    (640, 360, 8, 7, 3),
    (96, 64, 10, 4, 4),
    (250, 141, 8, 6, 5),
    (400, 300, 10, 2, 6),
]


def make_clip(w, h, bits, n, seed):
    c = synth.video_clip(w, h, bits, n, seed)
    pairs = [c.frame(f) for f in range(n)]
    fmt = api.PF_YUV420_U8 if bits == 8 else api.PF_YUV420_U16
    matrix = api.MATRIX_BT709 if bits == 8 else api.MATRIX_BT2020NCL
    return api.DevicePairs(pairs, fmt, bits, matrix)


@pytest.fixture(scope="module")
def clips():
    vship_amd.gpu_full_check(0)
    dps = [make_clip(*c) for c in CLIPS]
    # reference: each frame alone, default map tiling
    alone = []
    for dp in dps:
        with api.SSIMU2Batch(1, 640, 360) as b:
            s = []
            one = (api.Pair * 1)()
            for i in range(dp.n):
                one[0] = dp.descs[i]
                s.append(b.compute(one, 1)[0])
        alone.append(np.array(s))
    yield dps, alone
    for dp in dps:
        dp.free()


def test_server_scores_equal_alone(clips):
    dps, alone = clips
    budget = 2 * (vship_amd.pair_footprint(640, 360) + vship_amd.pair_footprint(400, 300))
    say(f"case: server, budget {budget} B, 4 pairs per step, 6 clips arriving in two waves")
    with api.SSIMU2Server(budget, 4, 640, 360, max_inflight_clips=3) as srv:
        ids = [srv.submit(dp.descs) for dp in dps[:3]]
        time.sleep(0.01)
        ids += [srv.submit(dp.descs) for dp in dps[3:]]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    for k, (scores, lat) in enumerate(res):
        say(f"clip {k} {CLIPS[k][:3]} latency {lat:.2f} ms")
        assert np.array_equal(scores, alone[k]), (k, scores, alone[k])
        assert lat > 0
    assert st["pairs_done"] == sum(c[3] for c in CLIPS)
    assert st["clips_done"] == len(CLIPS)


def test_server_poll_and_drain_on_free(clips):
    dps, alone = clips
    budget = 4 * vship_amd.pair_footprint(640, 360)
    srv = api.SSIMU2Server(budget, 8, 640, 360)
    t0 = srv.clock_ms()
    ids = [srv.submit(dp.descs, arrival_ms=t0) for dp in dps]
    got = {}
    deadline = time.time() + 60
    while len(got) < len(ids) and time.time() < deadline:
        for i in ids:
            if i not in got:
                r = srv.poll(i)
                if r is not None:
                    got[i] = r
        time.sleep(0.001)
    assert len(got) == len(ids)
    for k, i in enumerate(ids):
        assert np.array_equal(got[i][0], alone[k])
    # a clip still queued when the server is freed is drained first
    srv.submit(dps[2].descs)
    srv.close()


def test_server_rejects_clip_over_budget(clips):
    dps, _ = clips
    small = vship_amd.pair_footprint(160, 90) * 2  # two in-flight steps of one 160x90 pair
    with api.SSIMU2Server(small, 2, 640, 360) as srv:
        with pytest.raises(api.VshipError) as e:
            srv.submit(dps[2].descs)  # 640x360 does not fit
        assert e.value.code == 1
        sid = srv.submit(dps[0].descs)
        scores, _ = srv.wait(sid)
        assert np.all((scores > -1000) & (scores <= 100))
