"""Host-input batches (Vship_SSIMU2BatchSubmitHost) and the per-frame
handles that share their upload path, against the device-resident batch path:

* pageable planes (through the page-locked staging ring, default and tiny
  chunks with several copy threads) and page-locked planes (direct DMA) score
  bit for bit as the same pairs already in device memory, for 8- and 10-bit
  YUV 4:2:0 and padded-stride float RGB;
* two batches in flight: Wait returns them in submit order, a third Submit
  and a Compute while one is outstanding are refused, and the staging shape
  cannot change after the first host submit;
* Vship_ComputeSSIMU2Float gives the same score for pageable and page-locked
  planes.
"""
import numpy as np
import pytest

import vship_amd
from vship_amd import api, synth
from tests.gpu_helpers import rgb_pair, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu_ready():
    vship_amd.gpu_full_check(0)
    yield


def yuv_pairs(w, h, bits, n, seed):
    c = synth.video_clip(w, h, bits, n, seed)
    pairs = [c.frame(f) for f in range(n)]
    c.close()
    return pairs


def padded_rgb_pairs(w, h, n, pad, seed):
    """float RGB planes as views of rows padded by `pad` floats."""
    out = []
    for i in range(n):
        r, d = rgb_pair(w, h, api.PF_RGB_F32, seed + i)
        views = []
        for img in (r, d):
            v = []
            for p in img:
                buf = np.full((h, w + pad), 9.0, np.float32)
                buf[:, :w] = p
                v.append(buf[:, :w])
            views.append(v)
        out.append(tuple(views))
    return out


GROUPS = [  # (name, pairs, fmt, bits, matrix, max w, max h)
    ("1080p YUV420P8", lambda: yuv_pairs(1920, 1080, 8, 3, 61), api.PF_YUV420_U8, 8, api.MATRIX_BT709),
    ("640x360 YUV420P10", lambda: yuv_pairs(640, 360, 10, 4, 62), api.PF_YUV420_U16, 10, api.MATRIX_BT2020NCL),
    ("333x187 RGB f32, padded rows", lambda: padded_rgb_pairs(333, 187, 3, 21, 63), api.PF_RGB_F32, 0, 0),
]


@pytest.mark.parametrize("gi", range(len(GROUPS)), ids=[g[0] for g in GROUPS])
def test_host_batches_equal_device_batches(gi):
    name, make, fmt, bits, matrix = GROUPS[gi]
    pairs = make()
    n = len(pairs)
    w = max(p[0][0].shape[1] for p in pairs)
    h = max(p[0][0].shape[0] for p in pairs)
    say(f"case: {n} pairs {name}, host input pageable / page-locked / small staging chunks vs device input")
    dp = api.DevicePairs([([np.ascontiguousarray(q) for q in r], [np.ascontiguousarray(q) for q in d])
                          for r, d in pairs], fmt, bits, matrix)
    pageable = api.HostPairs(pairs, fmt, bits, matrix)
    pinned = api.HostPairs(pairs, fmt, bits, matrix, pinned=True)
    try:
        with api.SSIMU2Batch(n, w, h) as b:
            want = b.compute(dp.descs)
            for rep in range(2):
                b.submit_host(pageable.descs)
                assert np.array_equal(b.wait(), want), ("pageable", rep)
                b.submit_host(pinned.descs)
                assert np.array_equal(b.wait(), want), ("page-locked", rep)
        with api.SSIMU2Batch(n, w, h) as b:
            # 64 KiB chunks through 2 slots, copied by 3 threads: many refills per plane
            b.host_staging(64 << 10, 2, 3)
            b.submit_host(pageable.descs)
            assert np.array_equal(b.wait(), want), "small chunks"
            with pytest.raises(api.VshipError) as e:
                b.host_staging(8 << 20, 4, 4)
            assert e.value.code == 9  # BadHandler: the ring exists
    finally:
        pinned.free()
        pageable.free()
        dp.free()
    say(f"  scores {np.round(want, 4)}")


def test_two_batches_in_flight_fifo():
    a_pairs = yuv_pairs(1280, 720, 8, 5, 64)
    b_pairs = yuv_pairs(1280, 720, 8, 2, 65)
    say("case: two host batches of 5 and 2 1280x720 YUV420P8 pairs in flight, then device + host interleaved")
    da = api.DevicePairs(a_pairs, api.PF_YUV420_U8, 8, api.MATRIX_BT709)
    db = api.DevicePairs(b_pairs, api.PF_YUV420_U8, 8, api.MATRIX_BT709)
    ha = api.HostPairs(a_pairs, api.PF_YUV420_U8, 8, api.MATRIX_BT709)
    hb = api.HostPairs(b_pairs, api.PF_YUV420_U8, 8, api.MATRIX_BT709, pinned=True)
    try:
        with api.SSIMU2Batch(5, 1280, 720) as b:
            wa = b.compute(da.descs)
            wb = b.compute(db.descs)
            for rep in range(3):
                b.submit_host(ha.descs)
                b.submit_host(hb.descs)
                with pytest.raises(api.VshipError) as e:
                    b.submit(da.descs)
                assert e.value.code == 9
                assert np.array_equal(b.wait(), wa), rep
                with pytest.raises(api.VshipError) as e:
                    b.compute(da.descs)
                assert e.value.code == 9
                assert np.array_equal(b.wait(), wb), rep
            # device and host batches interleaved two deep
            b.submit(db.descs)
            b.submit_host(ha.descs)
            assert np.array_equal(b.wait(), wb)
            b.submit(da.descs)
            assert np.array_equal(b.wait(), wa)
            assert np.array_equal(b.wait(), wa)
            with pytest.raises(api.VshipError) as e:
                b.wait()
            assert e.value.code == 9
    finally:
        for x in (hb, ha, db, da):
            x.free()


def test_per_frame_handle_pageable_and_pinned_planes_agree():
    w, h = 1920, 1080
    r, d = rgb_pair(w, h, api.PF_RGB_F32, 66)
    say(f"case: Vship_ComputeSSIMU2Floatv2 {w}x{h}, pageable vs page-locked planes")
    pin = api.PinnedBuffer(6 * w * h * 4)
    try:
        views = []
        for i, p in enumerate(list(r) + list(d)):
            v = pin.array[i * w * h * 4:(i + 1) * w * h * 4].view(np.float32).reshape(h, w)
            v[...] = p
            views.append(v)
        with vship_amd.SSIMU2Handler(w, h) as hd:
            s_pageable = [hd.compute_float(r, d) for _ in range(2)]
            s_pinned = [hd.compute_float(views[:3], views[3:]) for _ in range(2)]
        say(f"  pageable {s_pageable} page-locked {s_pinned}")
        assert s_pageable[0] == s_pageable[1] == s_pinned[0] == s_pinned[1]
    finally:
        pin.free()


def test_mixed_page_locked_and_pageable_pieces():
    """One batch whose planes mix page-locked planes of one allocation (some at
    256-byte multiples, which one DMA carries together, some at 64-byte
    offsets, which it cannot), pageable planes, and a pair whose reference and
    distorted planes are the same page-locked memory."""
    w, h = 720, 405
    pairs = yuv_pairs(w, h, 8, 4, 67)
    say(f"case: 5 pairs {w}x{h} YUV420P8, page-locked / pageable / odd-offset / aliased planes in one batch")
    dp = api.DevicePairs(pairs + [(pairs[0][0], pairs[0][0])], api.PF_YUV420_U8, 8, api.MATRIX_BT709)
    sizes = [p.nbytes for r, d in pairs[:2] for p in list(r) + list(d)]
    pin = api.PinnedBuffer(sum(sizes) + 64 * 12 + 4096)
    pageable = api.HostPairs(pairs, api.PF_YUV420_U8, 8, api.MATRIX_BT709)
    try:
        descs = (api.Pair * 5)()
        for i in range(4):
            descs[i] = pageable.descs[i]
        # pairs 0 and 1 in page-locked memory: pair 0's planes at 256-byte
        # multiples, pair 1's at 64-byte offsets
        off = 0
        for i in (0, 1):
            planes = list(pairs[i][0]) + list(pairs[i][1])
            ptrs = []
            for p in planes:
                step = 256 if i == 0 else 64
                off = (off + step - 1) // step * step + (64 if i == 1 else 0)
                v = pin.array[off:off + p.nbytes].view(p.dtype).reshape(p.shape)
                v[...] = p
                ptrs.append(v.ctypes.data)
                off += p.nbytes
            for j in range(3):
                descs[i].ref[j] = ptrs[j]
                descs[i].dist[j] = ptrs[3 + j]
        # pair 4: pair 0's page-locked reference planes as both images
        descs[4] = descs[0]
        for j in range(3):
            descs[4].dist[j] = descs[0].ref[j]
        with api.SSIMU2Batch(5, w, h) as b:
            want = b.compute(dp.descs)
            for rep in range(2):
                b.submit_host(descs)
                got = b.wait()
                assert np.array_equal(got, want), (rep, got, want)
        assert want[4] == 100.0
    finally:
        pageable.free()
        pin.free()
        dp.free()
