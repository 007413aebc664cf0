"""Vship_SSIMU2ComputeNoAlloc, the allocation-free entry point (a fixed
scratch area per GPU instead of an engine's allocations), against
Vship_SSIMU2BatchCompute at its defaults on the same device-resident pairs:
two calls in a row, the second on a smaller pair (the scratch keeps the first
call's contents), scores and all 108 features bit-exact; sizes beyond the
scratch and host outputs are refused before anything runs."""
import ctypes

import numpy as np
import pytest

import vship_amd
from vship_amd import api, synth
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu

VSHIP_MAP_MFMA = 1


@pytest.fixture(scope="module", autouse=True)
def gpu_ready():
    vship_amd.gpu_full_check(0)
    yield


def batch_reference(dp):
    """Vship_SSIMU2BatchCompute with the default kernels: matrix-core map, one workgroup per tile."""
    w, h = dp.descs[0].width, dp.descs[0].height
    with api.SSIMU2Batch(1, w, h, map_tile_height=VSHIP_MAP_MFMA) as b:
        return b.compute(dp.descs, 1, features=True)


def noalloc(dp, out):
    """One call of the allocation-free entry point on the null stream; out is
    a device buffer of 109 doubles (score, then 108 features)."""
    rc = api.lib().Vship_SSIMU2ComputeNoAlloc(dp.descs, out.ptr, out.ptr + 8, None)
    assert rc == 0, rc
    raw = out.download(109 * 8)  # synchronous copy: waits for the null stream
    v = raw.view(np.float64)
    return v[0], v[1:].copy()


def test_no_alloc_matches_batch_entry_point():
    out = api.DeviceBuffer(109 * 8)
    try:
        # a large pair first, then a smaller one of another format: the second
        # call runs on the scratch the first one left
        cases = [(1920, 1080, 8, 21), (333, 187, 10, 22)]
        for w, h, bits, seed in cases:
            c = synth.video_clip(w, h, bits, 1, seed)
            r, d = c.frame(0)
            c.close()
            fmt = api.PF_YUV420_U8 if bits == 8 else api.PF_YUV420_U16
            matrix = api.MATRIX_BT709 if bits == 8 else api.MATRIX_BT2020NCL
            dp = api.DevicePairs([(r, d)], fmt, bits, matrix)
            try:
                say(f"case: no-alloc {w}x{h} {bits}-bit vs Vship_SSIMU2BatchCompute")
                s_na, f_na = noalloc(dp, out)
                (s_b,), (f_b,) = batch_reference(dp)
                say(f"  no-alloc {s_na!r} batch {s_b!r}")
                assert s_na == s_b
                assert np.array_equal(f_na, f_b)
            finally:
                dp.free()
    finally:
        out.free()


def test_no_alloc_refuses_before_running():
    out = api.DeviceBuffer(109 * 8)
    try:
        # 4K needs more than the scratch's arena: OutOfVRAM, nothing enqueued
        c = synth.video_clip(3840, 2160, 8, 1, 5)
        r, d = c.frame(0)
        c.close()
        dp = api.DevicePairs([(r, d)], api.PF_YUV420_U8, 8, api.MATRIX_BT709)
        try:
            say("case: no-alloc 3840x2160 refused (scratch too small)")
            assert api.lib().Vship_SSIMU2ComputeNoAlloc(dp.descs, out.ptr, None, None) == 1  # OutOfVRAM
            # a host score pointer is refused
            host = np.zeros(1)
            assert api.lib().Vship_SSIMU2ComputeNoAlloc(dp.descs, host.ctypes.data, None, None) == 10
        finally:
            dp.free()
    finally:
        out.free()
