"""CPU checks of the C-ABI library: it loads, exports every symbol the public
headers declare, mirrors the reference's error protocol and struct layouts.
No GPU work is issued here."""
import ctypes
import os
import re
import subprocess

import pytest

import vship_amd
from vship_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("VshipAPI.h", "VshipSSIMU2Batch.h", "VshipButteraugliBatch.h")]


def declared_symbols():
    names = []
    for h in HEADERS:
        txt = open(h).read()
        names += re.findall(r"EXPORTPREPROCESS\s+[\w\s\*]+?\b(Vship_\w+)\s*\(", txt)
    return sorted(set(names))


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", api.lib_path()], capture_output=True, text=True).stdout
    return {l.split()[-1] for l in out.splitlines() if l.strip()}


def test_library_exports_every_declared_symbol():
    decl = declared_symbols()
    assert len(decl) >= 42
    missing = [d for d in decl if d not in exported_symbols()]
    assert not missing, missing


def test_reference_symbols_present():
    # the reference's own ABI (src/VshipAPI.h:32-148) must all be there
    ref = ["Vship_GetVersion", "Vship_GetDeviceCount", "Vship_SetDevice", "Vship_GetDeviceInfo",
           "Vship_GPUFullCheck", "Vship_GetErrorMessage", "Vship_PinnedMalloc", "Vship_PinnedFree",
           "Vship_SSIMU2Init", "Vship_SSIMU2Free", "Vship_ComputeSSIMU2Float", "Vship_ComputeSSIMU2Floatv2",
           "Vship_ComputeSSIMU2Uint16", "Vship_ComputeSSIMU2Uint16v2", "Vship_ButteraugliInit",
           "Vship_ButteraugliInitv2", "Vship_ButteraugliFree", "Vship_ComputeButteraugliFloat",
           "Vship_ComputeButteraugliFloatv2", "Vship_ComputeButteraugliUint16", "Vship_ComputeButteraugliUint16v2"]
    exp = exported_symbols()
    assert [r for r in ref if r not in exp] == []


def test_version():
    assert vship_amd.version() == (3, 1, 0, 0)


def test_error_message_protocol():
    L = api.lib()
    for code in range(12):
        n = L.Vship_GetErrorMessage(code, None, 0)
        msg = api.error_message(code)
        assert n == len(msg) + 1
        assert msg.startswith(api.ERRORS[code] + ":")
        # truncation: len - 1 characters then NUL
        buf = ctypes.create_string_buffer(b"\xff" * 8)
        assert L.Vship_GetErrorMessage(code, buf, 8) == n
        assert buf.raw[7] == 0 and buf.value == msg[:7].encode()


def test_error_messages_are_the_reference_texts():
    assert api.error_message(9).startswith("BadHandler: The handler used was not allocated.")
    assert api.error_message(1) == ("OutOfVRAM: Vship was not able to perform GPU memory allocation. "
                                    "(Advice) Reduce or Set numStream argument")


def test_free_of_unknown_handler_is_bad_handler():
    L = api.lib()
    assert L.Vship_SSIMU2Free(api.Handler(123456)) == 9
    assert L.Vship_SSIMU2Free(api.Handler(-1)) == 9
    assert L.Vship_SSIMU2BatchFree(api.Handler(777)) == 9


def test_compute_on_unknown_handler_is_bad_handler():
    L = api.lib()
    score = ctypes.c_double(-5.0)
    p = api.P3(0, 0, 0)
    assert L.Vship_ComputeSSIMU2Float(api.Handler(4242), ctypes.byref(score), p, p, 16) == 9
    assert score.value == -5.0  # written only on success


def test_butteraugli_error_protocol_without_gpu():
    # src/VshipLib.cpp:252-296: the slot is taken even when init fails, a bad
    # handler is reported as BadHandler, and nothing is written on failure
    L = api.lib()
    h = api.Handler()
    code = L.Vship_ButteraugliInit(ctypes.byref(h), 16, 16, ctypes.c_float(80.0))
    assert code != 0
    assert L.Vship_ButteraugliFree(h) == 0
    score = api.ButterScore(norm2=-7, norm3=-7, norminf=-7)
    P3 = ctypes.c_void_p * 3
    assert L.Vship_ComputeButteraugliFloat(api.Handler(id=999), ctypes.byref(score), None, 0, P3(), P3(), 64) == 9
    assert score.norm2 == -7
    assert L.Vship_ButteraugliPairFootprint(0, 5) == 0
    assert L.Vship_ButteraugliPairFootprint(64, 48) > 0
    cfg = api.ButterBatchConfig(width=16, height=16, max_pairs=0, qnorm=2, intensity_multiplier=203.0)
    assert L.Vship_ButteraugliBatchInit(ctypes.byref(h), ctypes.byref(cfg)) == 10


def test_struct_layouts_match_c(tmp_path):
    exe = tmp_path / "abi_layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "abi_layout.c"),
                    "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        k, *v = line.split()
        got[k] = [int(x) for x in v]
    assert got["Vship_Version"] == [ctypes.sizeof(api.Version)]
    assert got["Vship_DeviceInfo"] == [ctypes.sizeof(api.DeviceInfo), api.DeviceInfo.VRAMSize.offset,
                                       api.DeviceInfo.WarpSize.offset]
    assert got["Vship_SSIMU2Pair"] == [ctypes.sizeof(api.Pair), api.Pair.width.offset, api.Pair.matrix.offset]
    assert got["Vship_SSIMU2BatchConfig"] == [ctypes.sizeof(api.BatchConfig), api.BatchConfig.arena_bytes.offset,
                                              api.BatchConfig.profile.offset]
    assert got["Vship_SSIMU2BatchStats"] == [ctypes.sizeof(api.BatchStats), api.BatchStats.map_bytes.offset]
    assert got["Vship_ButteraugliScore"] == [12]
    assert got["Vship_ButteraugliBatchConfig"] == [ctypes.sizeof(api.ButterBatchConfig),
                                                   api.ButterBatchConfig.intensity_multiplier.offset]
    assert got["Vship_ButteraugliPair"] == [ctypes.sizeof(api.ButterPair), api.ButterPair.distmap.offset]
    assert got["Vship_ButteraugliBatchStats"] == [ctypes.sizeof(api.ButterBatchStats),
                                                  api.ButterBatchStats.stage_bytes.offset]


def test_device_count_without_gpu_reports_an_error_code():
    import shutil
    if os.path.exists("/dev/kfd") and shutil.which("rocminfo") and os.environ.get("VSHIP_EXPECT_GPU"):
        pytest.skip("GPU present")
    n = ctypes.c_int(-1)
    code = api.lib().Vship_GetDeviceCount(ctypes.byref(n))
    if code == 0:
        assert n.value >= 1
    else:
        assert code in (5, 6)


def test_pair_footprint_matches_layout():
    # 6 scales x 3 planes x 2 images, each plane row padded to 64 elements
    # (FP16 storage: scales 0 and 1 in 2 bytes, scales 2-5 in 4)
    w, h = 1920, 1080
    tot = half = 0
    for s in range(6):
        plane = ((w + 63) // 64 * 64) * h
        tot += plane * 4
        half += plane * (2 if s < 2 else 4)
        w, h = (w - 1) // 2 + 1, (h - 1) // 2 + 1
    assert vship_amd.pair_footprint(1920, 1080) == tot * 3 * 2
    assert vship_amd.pair_footprint(1920, 1080, fp16=True) == half * 3 * 2
    assert vship_amd.pair_footprint(0, 10) == 0
