"""The VapourSynth plugin (vship.SSIMULACRA2, vship.BUTTERAUGLI, vship.GpuInfo) through a mock
VapourSynth host (tests/vs_mock_host.cpp): registration, argument and error
handling on the CPU; scores, output frames and reference counting on the GPU.
The plugin's own declarations of the VapourSynth API (vship_amd/csrc/vs4.hpp)
are pinned against VapourSynth's VapourSynth4.h where that header is present.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from vship_amd import api

from tests.vs_mock import MOCK, VS_HEADER_DIR, libpath, mock, run_butter, run_ssimu2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

needs_mock = pytest.mark.skipif(not os.path.exists(MOCK), reason="mock host not built (python -m vship_amd.build)")


def padded_frames(n, w, h, stride_px, seed):
    """n pairs of planar float RGB in [0, 1] with rows padded to stride_px floats."""
    from tests.gpu_helpers import rgb_pair

    ref, dist = [], []
    for i in range(n):
        r, d = rgb_pair(w, h, api.PF_RGB_F32, seed + i)
        pr, pd = [], []
        for a, b in zip(r, d):
            x = np.zeros((h, stride_px), np.float32)
            y = np.full((h, stride_px), 7.0, np.float32)  # padding that must not be read
            x[:, :w] = a
            y[:, :w] = b
            pr.append(x)
            pd.append(y)
        ref.append(pr)
        dist.append(pd)
    return ref, dist


# ---------------------------------------------------------------- CPU
@pytest.mark.skipif(not os.path.exists(os.path.join(VS_HEADER_DIR, "VapourSynth4.h")) or not shutil.which("g++"),
                    reason="VapourSynth4.h not available")
def test_vs4_declarations_match_vapoursynth_header(tmp_path):
    exe = str(tmp_path / "vs4chk")
    subprocess.run(["g++", "-std=c++17", "-I", VS_HEADER_DIR, os.path.join(ROOT, "tests", "vs4_layout_check.cpp"),
                    "-o", exe], check=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 0, r.stdout
    assert "0 mismatches" in r.stdout


@needs_mock
def test_plugin_registers_the_reference_interface():
    buf = ctypes.create_string_buffer(8192)
    assert mock().vsmock_registration(libpath(), buf, len(buf)) == 0
    lines = buf.value.decode().splitlines()
    # src/VshipLib.cpp:59-63 of the reference
    assert lines[0] == f"config|com.lumen.vship|vship|VapourSynth GPU Perceptual Metrics|{(3 << 16) | 3}|{4 << 16}"
    fns = {l.split("|")[1]: l.split("|")[2:] for l in lines[1:]}
    assert fns["SSIMULACRA2"] == ["reference:vnode;distorted:vnode;numStream:int:opt;gpu_id:int:opt;", "clip:vnode;"]
    assert fns["GpuInfo"] == ["gpu_id:int:opt;", "gpu_human_data:data;"]
    # src/butter/vapoursynth.hpp argument names, src/VshipLib.cpp:61
    assert fns["BUTTERAUGLI"] == ["reference:vnode;distorted:vnode;qnorm:int:opt;intensity_multiplier:float:opt;"
                                  "distmap:int:opt;numStream:int:opt;gpu_id:int:opt;", "clip:vnode;"]
    assert set(fns) == {"SSIMULACRA2", "GpuInfo", "BUTTERAUGLI", "CVVDP"}


@needs_mock
def test_differing_inputs_are_rejected_before_any_gpu_work():
    ref, dist = padded_frames(1, 64, 48, 64, 1)
    rc, _, _, err = run_ssimu2(ref, dist, 64, 48, 64 * 4, w2=32)
    assert rc == 1, err
    assert err.startswith("VshipException\nDifferingInputType: Vship received 2 videos with different properties")
    assert "\n - At line " in err and err.endswith("\n")


@needs_mock
def test_butteraugli_differing_inputs_are_rejected():
    ref, dist = padded_frames(1, 64, 48, 64, 1)
    rc, _, _, _, err = run_butter(ref, dist, 64, 48, 64 * 4, h2=40)
    assert rc == 1, err
    assert err.startswith("VshipException\nDifferingInputType:")


@needs_mock
def test_non_rgbs_input_without_resize_plugin_is_an_error():
    ref, dist = padded_frames(1, 64, 48, 64, 1)
    rc, _, _, err = run_ssimu2(ref, dist, 64, 48, 64 * 4, rgbs=(0, 1))
    assert rc == 1, err  # 3 would mean a node reference leaked or was freed twice
    assert err.startswith("VshipException\nNonRGBSInput:")


@needs_mock
def test_no_gpu_is_reported_in_the_reference_format():
    import vship_amd

    try:
        if vship_amd.device_count() > 0:
            pytest.skip("a GPU is present")
    except api.VshipError:
        pass  # no device: the case under test
    ref, dist = padded_frames(1, 64, 48, 64, 1)
    rc, _, _, err = run_ssimu2(ref, dist, 64, 48, 64 * 4)
    assert rc == 1
    assert err.startswith("VshipException\nDeviceCountError:") or err.startswith("VshipException\nNoDeviceDetected:")
    buf = ctypes.create_string_buffer(4096)
    assert mock().vsmock_gpuinfo(libpath(), -1, buf, len(buf)) == 1
    assert buf.value.decode().startswith("VshipException\n")


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
@needs_mock
def test_ssimulacra2_filter_scores_every_frame():
    from tests import oracle_lib as ol
    import vship_amd

    vship_amd.gpu_full_check(0)
    w, h, stride_px, n = 320, 240, 336, 6
    ref, dist = padded_frames(n, w, h, stride_px, 50)
    print(f"case: vship.SSIMULACRA2 via the mock host, {n} frames {w}x{h}, numStream 3, 4 threads", flush=True)
    rc, scores, matches, err = run_ssimu2(ref, dist, w, h, stride_px * 4, num_stream=3, gpu_id=0, nthreads=4)
    assert rc == 0, err
    assert matches == n  # output frame = distorted frame, frame n requests frame n of both clips
    with vship_amd.SSIMU2Handler(w, h) as hd:
        direct = np.array([hd.compute_float([p[:, :w] for p in r], [p[:, :w] for p in d]) for r, d in zip(ref, dist)])
    oracle = np.array([ol.score_rgb(ol.SO_RGB_F32, [p[:, :w] for p in r], [p[:, :w] for p in d], "f64")[0]
                       for r, d in zip(ref, dist)])
    print(f"filter {scores}\noracle {oracle}", flush=True)
    assert np.array_equal(scores, direct)
    assert np.all(np.abs(scores - oracle) <= 1e-3)


@pytest.mark.gpu
@needs_mock
def test_gpuinfo():
    buf = ctypes.create_string_buffer(8192)
    assert mock().vsmock_gpuinfo(libpath(), -1, buf, len(buf)) == 0
    assert buf.value.decode().startswith("GPU 0: ")
    assert mock().vsmock_gpuinfo(libpath(), 0, buf, len(buf)) == 0
    text = buf.value.decode()
    assert "WarpSize: 64" in text and "PassKernelCheck : 1" in text
    assert mock().vsmock_gpuinfo(libpath(), 99, buf, len(buf)) == 1
    assert buf.value.decode().startswith("VshipException\nBadDeviceArgument:")


@pytest.mark.gpu
@needs_mock
def test_butteraugli_filter_scores_every_frame():
    from tests import butter_lib as bl
    import vship_amd

    vship_amd.gpu_full_check(0)
    w, h, stride_px, n = 200, 120, 208, 4
    ref, dist = padded_frames(n, w, h, stride_px, 70)
    print(f"case: vship.BUTTERAUGLI via the mock host, {n} frames {w}x{h}, qnorm 3, distmap", flush=True)
    rc, norms, maps, _, err = run_butter(ref, dist, w, h, stride_px * 4, qnorm=3, distmap=1, num_stream=2, gpu_id=0)
    assert rc == 0, err
    for i in range(n):
        r = [np.ascontiguousarray(p[:, :w]) for p in ref[i]]
        d = [np.ascontiguousarray(p[:, :w]) for p in dist[i]]
        want, wmap = bl.butteraugli(r, d, 3, 203.0, diffmap=True)
        assert np.all(np.abs(norms[i] - want) <= 1e-4 * want), (norms[i], want)
        assert np.all(np.abs(maps[i] - wmap) <= 1e-6 + 1e-4 * np.abs(wmap))
    print("case: vship.BUTTERAUGLI defaults, output = distorted frames", flush=True)
    rc, norms2, _, matches, err = run_butter(ref, dist, w, h, stride_px * 4, gpu_id=0)
    assert rc == 0, err
    assert matches == n
    for i in range(n):
        r = [np.ascontiguousarray(p[:, :w]) for p in ref[i]]
        d = [np.ascontiguousarray(p[:, :w]) for p in dist[i]]
        want = bl.butteraugli(r, d, 2, 203.0)
        assert np.all(np.abs(norms2[i] - want) <= 1e-4 * want)



@pytest.mark.gpu
@needs_mock
def test_ssimulacra2_filter_1080p_throughput():
    """48 frames of 1920x1080 float RGB in pageable memory (16 distinct
    pairs, each requested three times) through the filter: every score equal
    to the per-frame entry point's, and the frame rate printed for 1, 3 and 8
    engines on 8 request threads."""
    import time

    import vship_amd

    vship_amd.gpu_full_check(0)
    w, h, uniq, reps = 1920, 1080, 16, 3
    t0 = time.time()
    from tests.gpu_helpers import rgb_pair

    pairs = [rgb_pair(w, h, api.PF_RGB_F32, 700 + i) for i in range(uniq)]
    print(f"case: vship.SSIMULACRA2 via the mock host, {uniq * reps} frames {w}x{h} float RGB "
          f"(generated in {time.time() - t0:.1f}s)", flush=True)
    ref = [pairs[i % uniq][0] for i in range(uniq * reps)]
    dist = [pairs[i % uniq][1] for i in range(uniq * reps)]
    with vship_amd.SSIMU2Handler(w, h) as hd:
        direct = np.array([hd.compute_float(r, d) for r, d in pairs])
    for ns in (1, 3, 8):
        rc, scores, matches, err, ms = run_ssimu2(ref, dist, w, h, w * 4, num_stream=ns, gpu_id=0, core_threads=8,
                                                  nthreads=8, timed=True)
        assert rc == 0, err
        assert matches == uniq * reps
        print(f"  numStream {ns}, 8 threads: {uniq * reps / ms * 1e3:.0f} pairs/s ({ms:.1f} ms)", flush=True)
        assert np.array_equal(scores, np.tile(direct, reps))
