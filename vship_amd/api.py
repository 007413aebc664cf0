"""ctypes binding of libvship.so (see the package docstring)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def lib_path():
    return os.environ.get("VSHIP_LIB", os.path.join(_HERE, "libvship.so"))


# Vship_Exception values, src/VshipAPI.h:35-58 of the reference
ERRORS = [
    "NoError", "OutOfVRAM", "OutOfRAM", "DifferingInputType", "NonRGBSInput", "DeviceCountError",
    "NoDeviceDetected", "BadDeviceArgument", "BadDeviceCode", "BadHandler", "BadPointer", "BadErrorType",
]

# Vship_PixelFormat / Vship_Matrix / Vship_ChromaLocation (include/VshipSSIMU2Batch.h)
(PF_RGB_U8, PF_RGB_U16, PF_RGB_F16, PF_RGB_F32, PF_YUV420_U8, PF_YUV420_U16, PF_YUV422_U8, PF_YUV422_U16,
 PF_YUV444_U8, PF_YUV444_U16, PF_YUV440_U8, PF_YUV440_U16, PF_YUV411_U8, PF_YUV411_U16, PF_YUV410_U8,
 PF_YUV410_U16) = range(16)
MATRIX_BT709, MATRIX_BT601, MATRIX_BT2020NCL = range(3)
CHROMA_LEFT, CHROMA_CENTER, CHROMA_TOPLEFT, CHROMA_TOP = range(4)
# YUV layout name -> (8-bit format, 16-bit-container format, log2 horizontal, log2 vertical subsampling)
YUV_LAYOUTS = {"420": (PF_YUV420_U8, PF_YUV420_U16, 1, 1), "422": (PF_YUV422_U8, PF_YUV422_U16, 1, 0),
               "444": (PF_YUV444_U8, PF_YUV444_U16, 0, 0), "440": (PF_YUV440_U8, PF_YUV440_U16, 0, 1),
               "411": (PF_YUV411_U8, PF_YUV411_U16, 2, 0), "410": (PF_YUV410_U8, PF_YUV410_U16, 2, 1)}


class VshipNotBuilt(RuntimeError):
    pass


class VshipError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = int(code)
        self.name = ERRORS[self.code] if 0 <= self.code < len(ERRORS) else f"code {self.code}"
        super().__init__(f"{self.name}: {error_message(self.code)} {where}".strip())


class Version(ctypes.Structure):
    _fields_ = [("major", ctypes.c_int), ("minor", ctypes.c_int), ("minorMinor", ctypes.c_int),
                ("backend", ctypes.c_int)]


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 256), ("VRAMSize", ctypes.c_uint64), ("integrated", ctypes.c_int),
                ("MultiProcessorCount", ctypes.c_int), ("WarpSize", ctypes.c_int)]


class Handler(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int)]


class Pair(ctypes.Structure):
    _fields_ = [("ref", ctypes.c_void_p * 3), ("dist", ctypes.c_void_p * 3),
                ("ref_stride", ctypes.c_int64 * 3), ("dist_stride", ctypes.c_int64 * 3),
                ("width", ctypes.c_int), ("height", ctypes.c_int), ("format", ctypes.c_int),
                ("bits", ctypes.c_int), ("matrix", ctypes.c_int), ("chroma_location", ctypes.c_int)]


class BatchConfig(ctypes.Structure):
    _fields_ = [("max_pairs", ctypes.c_int), ("max_width", ctypes.c_int), ("max_height", ctypes.c_int),
                ("fp16_planes", ctypes.c_int), ("arena_bytes", ctypes.c_uint64),
                ("map_tile_height", ctypes.c_int), ("persistent_grid", ctypes.c_int),
                ("use_graph", ctypes.c_int), ("profile", ctypes.c_int)]


class BatchStats(ctypes.Structure):
    _fields_ = [("prep_ms", ctypes.c_float), ("map_ms", ctypes.c_float), ("map_fp32_ms", ctypes.c_float),
                ("map_fp64_ms", ctypes.c_float), ("finalize_ms", ctypes.c_float), ("total_ms", ctypes.c_float),
                ("npairs", ctypes.c_int), ("prep_tiles", ctypes.c_int), ("map_tiles", ctypes.c_int),
                ("map_tiles_fp32", ctypes.c_int), ("map_tiles_fp64", ctypes.c_int),
                ("map_tile_height", ctypes.c_int), ("map_bytes", ctypes.c_uint64),
                ("map_bytes_fp32", ctypes.c_uint64), ("prep_bytes", ctypes.c_uint64)]


class ButterScore(ctypes.Structure):
    _fields_ = [("norm2", ctypes.c_int), ("norm3", ctypes.c_int), ("norminf", ctypes.c_int)]


class ButterBatchConfig(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("max_pairs", ctypes.c_int),
                ("qnorm", ctypes.c_int), ("intensity_multiplier", ctypes.c_float), ("profile", ctypes.c_int)]


class ButterPair(ctypes.Structure):
    _fields_ = [("ref", ctypes.c_void_p * 3), ("dist", ctypes.c_void_p * 3),
                ("ref_stride", ctypes.c_int64 * 3), ("dist_stride", ctypes.c_int64 * 3),
                ("distmap", ctypes.c_void_p), ("distmap_stride", ctypes.c_int64), ("bits", ctypes.c_int),
                ("matrix", ctypes.c_int), ("chroma_location", ctypes.c_int)]


class ButterBatchStats(ctypes.Structure):
    _fields_ = [("input_ms", ctypes.c_float), ("opsin_ms", ctypes.c_float), ("low_ms", ctypes.c_float),
                ("mid_ms", ctypes.c_float), ("high_ms", ctypes.c_float), ("diff_ms", ctypes.c_float),
                ("mask_ms", ctypes.c_float), ("norm_ms", ctypes.c_float), ("npairs", ctypes.c_int),
                ("timed", ctypes.c_int), ("stage_bytes", ctypes.c_uint64 * 8)]


BUTTER_STAGES = ["input", "opsin", "low", "mid", "high", "diff", "mask", "norm"]


class TuneCandidate(ctypes.Structure):
    _fields_ = [("map_tile_height", ctypes.c_int), ("persistent_grid", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("ms", ctypes.c_float)]


class TuneResult(ctypes.Structure):
    _fields_ = [("map_tile_height", ctypes.c_int), ("persistent_grid", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("best_ms", ctypes.c_float), ("ncandidates", ctypes.c_int), ("candidates", TuneCandidate * 32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("memory_budget", ctypes.c_uint64), ("max_pairs_per_step", ctypes.c_int),
                ("max_inflight_clips", ctypes.c_int), ("max_width", ctypes.c_int), ("max_height", ctypes.c_int),
                ("fp16_planes", ctypes.c_int), ("map_tile_height", ctypes.c_int), ("persistent_grid", ctypes.c_int),
                ("use_graph", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [("pairs_done", ctypes.c_int64), ("steps", ctypes.c_int64), ("clips_done", ctypes.c_int64),
                ("host_ms", ctypes.c_double), ("waiting_clips", ctypes.c_int), ("active_clips", ctypes.c_int)]


_lib = None
P3 = ctypes.c_void_p * 3


def lib():
    """The loaded native library; raises VshipNotBuilt when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise VshipNotBuilt(f"{path} is not built: run `python -m vship_amd.build`")
    L = ctypes.CDLL(path)
    E = ctypes.c_int
    L.Vship_GetVersion.restype = Version
    L.Vship_GetDeviceCount.restype = E
    L.Vship_GetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
    L.Vship_SetDevice.restype = E
    L.Vship_SetDevice.argtypes = [ctypes.c_int]
    L.Vship_GetDeviceInfo.restype = E
    L.Vship_GetDeviceInfo.argtypes = [ctypes.POINTER(DeviceInfo), ctypes.c_int]
    L.Vship_GPUFullCheck.restype = E
    L.Vship_GPUFullCheck.argtypes = [ctypes.c_int]
    L.Vship_GetErrorMessage.restype = ctypes.c_int
    L.Vship_GetErrorMessage.argtypes = [E, ctypes.c_char_p, ctypes.c_int]
    L.Vship_PinnedMalloc.restype = E
    L.Vship_PinnedMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint64]
    L.Vship_PinnedFree.restype = E
    L.Vship_PinnedFree.argtypes = [ctypes.c_void_p]
    L.Vship_SSIMU2Init.restype = E
    L.Vship_SSIMU2Init.argtypes = [ctypes.POINTER(Handler), ctypes.c_int, ctypes.c_int]
    L.Vship_SSIMU2Free.restype = E
    L.Vship_SSIMU2Free.argtypes = [Handler]
    for name in ("Vship_ComputeSSIMU2Float", "Vship_ComputeSSIMU2Uint16"):
        fn = getattr(L, name)
        fn.restype = E
        fn.argtypes = [Handler, ctypes.POINTER(ctypes.c_double), P3, P3, ctypes.c_int64]
    for name in ("Vship_ComputeSSIMU2Floatv2", "Vship_ComputeSSIMU2Uint16v2"):
        fn = getattr(L, name)
        fn.restype = E
        fn.argtypes = [Handler, ctypes.POINTER(ctypes.c_double), P3, P3, ctypes.c_int64, ctypes.c_int64]
    L.Vship_SSIMU2PairFootprint.restype = ctypes.c_uint64
    L.Vship_SSIMU2PairFootprint.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.Vship_SSIMU2BatchInit.restype = E
    L.Vship_SSIMU2BatchInit.argtypes = [ctypes.POINTER(Handler), ctypes.POINTER(BatchConfig)]
    L.Vship_SSIMU2BatchFree.restype = E
    L.Vship_SSIMU2BatchFree.argtypes = [Handler]
    L.Vship_SSIMU2BatchSubmit.restype = E
    L.Vship_SSIMU2BatchSubmit.argtypes = [Handler, ctypes.POINTER(Pair), ctypes.c_int]
    L.Vship_SSIMU2BatchSubmitHost.restype = E
    L.Vship_SSIMU2BatchSubmitHost.argtypes = [Handler, ctypes.POINTER(Pair), ctypes.c_int]
    L.Vship_SSIMU2BatchHostStaging.restype = E
    L.Vship_SSIMU2BatchHostStaging.argtypes = [Handler, ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
    L.Vship_SSIMU2BatchWait.restype = E
    L.Vship_SSIMU2BatchWait.argtypes = [Handler, ctypes.c_void_p, ctypes.c_void_p]
    L.Vship_SSIMU2BatchCompute.restype = E
    L.Vship_SSIMU2BatchCompute.argtypes = [Handler, ctypes.POINTER(Pair), ctypes.c_int, ctypes.c_void_p,
                                           ctypes.c_void_p]
    L.Vship_SSIMU2BatchGetStats.restype = E
    L.Vship_SSIMU2BatchGetStats.argtypes = [Handler, ctypes.POINTER(BatchStats)]
    L.Vship_DeviceMalloc.restype = E
    L.Vship_DeviceMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint64]
    L.Vship_DeviceFree.restype = E
    L.Vship_DeviceFree.argtypes = [ctypes.c_void_p]
    L.Vship_MemcpyHtoD.restype = E
    L.Vship_MemcpyHtoD.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
    L.Vship_MemcpyDtoH.restype = E
    L.Vship_MemcpyDtoH.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
    L.Vship_MemcpyDtoD.restype = E
    L.Vship_MemcpyDtoD.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
    L.Vship_SSIMU2ComputeNoAlloc.restype = E
    L.Vship_SSIMU2ComputeNoAlloc.argtypes = [ctypes.POINTER(Pair), ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p]
    L.Vship_SSIMU2ServerInit.restype = E
    L.Vship_SSIMU2ServerInit.argtypes = [ctypes.POINTER(Handler), ctypes.POINTER(ServerConfig)]
    L.Vship_SSIMU2ServerFree.restype = E
    L.Vship_SSIMU2ServerFree.argtypes = [Handler]
    L.Vship_SSIMU2ServerSubmitClip.restype = E
    L.Vship_SSIMU2ServerSubmitClip.argtypes = [Handler, ctypes.POINTER(Pair), ctypes.c_int, ctypes.c_double,
                                              ctypes.POINTER(ctypes.c_int)]
    L.Vship_SSIMU2ServerPollClip.restype = E
    L.Vship_SSIMU2ServerPollClip.argtypes = [Handler, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_void_p,
                                            ctypes.POINTER(ctypes.c_double)]
    L.Vship_SSIMU2ServerWaitClip.restype = E
    L.Vship_SSIMU2ServerWaitClip.argtypes = [Handler, ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
    L.Vship_SSIMU2ServerGetStats.restype = E
    L.Vship_SSIMU2ServerGetStats.argtypes = [Handler, ctypes.POINTER(ServerStats)]
    L.Vship_SSIMU2BatchTune.restype = E
    L.Vship_SSIMU2BatchTune.argtypes = [ctypes.POINTER(Pair), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(TuneResult)]
    L.Vship_SSIMU2TuneSave.restype = E
    L.Vship_SSIMU2TuneSave.argtypes = [ctypes.c_char_p]
    L.Vship_SSIMU2TuneLoad.restype = E
    L.Vship_SSIMU2TuneLoad.argtypes = [ctypes.c_char_p]
    L.Vship_ButteraugliInit.restype = E
    L.Vship_ButteraugliInit.argtypes = [ctypes.POINTER(Handler), ctypes.c_int, ctypes.c_int, ctypes.c_float]
    L.Vship_ButteraugliInitv2.restype = E
    L.Vship_ButteraugliInitv2.argtypes = [ctypes.POINTER(Handler), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_float]
    L.Vship_ButteraugliFree.restype = E
    L.Vship_ButteraugliFree.argtypes = [Handler]
    for name in ("Vship_ComputeButteraugliFloat", "Vship_ComputeButteraugliUint16"):
        fn = getattr(L, name)
        fn.restype = E
        fn.argtypes = [Handler, ctypes.POINTER(ButterScore), ctypes.c_void_p, ctypes.c_int64, P3, P3, ctypes.c_int64]
    for name in ("Vship_ComputeButteraugliFloatv2", "Vship_ComputeButteraugliUint16v2"):
        fn = getattr(L, name)
        fn.restype = E
        fn.argtypes = [Handler, ctypes.POINTER(ButterScore), ctypes.c_void_p, ctypes.c_int64, P3, P3, ctypes.c_int64,
                       ctypes.c_int64]
    L.Vship_ButteraugliPairFootprint.restype = ctypes.c_uint64
    L.Vship_ButteraugliPairFootprint.argtypes = [ctypes.c_int, ctypes.c_int]
    L.Vship_ButteraugliBatchInit.restype = E
    L.Vship_ButteraugliBatchInit.argtypes = [ctypes.POINTER(Handler), ctypes.POINTER(ButterBatchConfig)]
    L.Vship_ButteraugliBatchFree.restype = E
    L.Vship_ButteraugliBatchFree.argtypes = [Handler]
    L.Vship_ButteraugliBatchCompute.restype = E
    L.Vship_ButteraugliBatchCompute.argtypes = [Handler, ctypes.c_int, ctypes.POINTER(ButterPair), ctypes.c_int,
                                                ctypes.c_void_p]
    L.Vship_ButteraugliBatchGetStats.restype = E
    L.Vship_ButteraugliBatchGetStats.argtypes = [Handler, ctypes.POINTER(ButterBatchStats)]
    L.Vship_SSIMU2ServerClockMs.restype = ctypes.c_double
    L.Vship_SSIMU2ServerClockMs.argtypes = []
    _lib = L
    return L


def _check(code, where=""):
    if code != 0:
        raise VshipError(code, where)


def error_message(code):
    L = lib()
    n = L.Vship_GetErrorMessage(int(code), None, 0)
    buf = ctypes.create_string_buffer(n)
    L.Vship_GetErrorMessage(int(code), buf, n)
    return buf.value.decode()


def version():
    v = lib().Vship_GetVersion()
    return (v.major, v.minor, v.minorMinor, v.backend)


def device_count():
    n = ctypes.c_int()
    _check(lib().Vship_GetDeviceCount(ctypes.byref(n)), "Vship_GetDeviceCount")
    return n.value


def set_device(i):
    _check(lib().Vship_SetDevice(int(i)), "Vship_SetDevice")


def device_info(i=0):
    d = DeviceInfo()
    _check(lib().Vship_GetDeviceInfo(ctypes.byref(d), int(i)), "Vship_GetDeviceInfo")
    return {"name": d.name.decode(errors="replace"), "VRAMSize": d.VRAMSize, "integrated": d.integrated,
            "MultiProcessorCount": d.MultiProcessorCount, "WarpSize": d.WarpSize}


def gpu_full_check(i=0):
    _check(lib().Vship_GPUFullCheck(int(i)), "Vship_GPUFullCheck")


def pair_footprint(w, h, fp16=False):
    return int(lib().Vship_SSIMU2PairFootprint(int(w), int(h), int(bool(fp16))))


class DeviceBuffer:
    """A device allocation made through Vship_DeviceMalloc."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib().Vship_DeviceMalloc(ctypes.byref(p), max(1, self.nbytes)), "Vship_DeviceMalloc")
        self.ptr = p.value

    def upload(self, arr, offset=0):
        a = np.ascontiguousarray(arr)
        if offset + a.nbytes > self.nbytes:
            raise ValueError("upload past the end of the buffer")
        _check(lib().Vship_MemcpyHtoD(self.ptr + offset, a.ctypes.data, a.nbytes), "Vship_MemcpyHtoD")

    def download(self, nbytes=None, offset=0):
        n = self.nbytes - offset if nbytes is None else nbytes
        out = np.empty(n, np.uint8)
        _check(lib().Vship_MemcpyDtoH(out.ctypes.data, self.ptr + offset, n), "Vship_MemcpyDtoH")
        return out

    def copy_from(self, other, nbytes):
        """Device-to-device copy of nbytes from another DeviceBuffer."""
        _check(lib().Vship_MemcpyDtoD(self.ptr, other.ptr, int(nbytes)), "Vship_MemcpyDtoD")

    def free(self):
        if self.ptr:
            lib().Vship_DeviceFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedBuffer:
    """Page-locked host memory through Vship_PinnedMalloc, viewed as numpy."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        _check(lib().Vship_PinnedMalloc(ctypes.byref(p), int(nbytes)), "Vship_PinnedMalloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)
        self.array = np.ctypeslib.as_array((ctypes.c_uint8 * self.nbytes).from_address(self.ptr))

    def free(self):
        if self.ptr:
            self.array = None
            _check(lib().Vship_PinnedFree(self.ptr), "Vship_PinnedFree")
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _plane_ptrs(planes):
    return P3(*[p.ctypes.data for p in planes])


class SSIMU2Handler:
    """Per-frame SSIMULACRA2 handle of fixed size: the reference's
    Vship_SSIMU2Init / Vship_ComputeSSIMU2{Float,Uint16}[v2] / Vship_SSIMU2Free
    (src/VshipAPI.h:87-111, src/VshipLib.cpp:162-250)."""

    def __init__(self, width, height):
        self.width, self.height = int(width), int(height)
        self._h = Handler()
        code = lib().Vship_SSIMU2Init(ctypes.byref(self._h), self.width, self.height)
        if code != 0:
            lib().Vship_SSIMU2Free(self._h)
            raise VshipError(code, "Vship_SSIMU2Init")

    @property
    def id(self):
        return self._h.id

    def _compute(self, fn, ref, dist, stride, stride2):
        score = ctypes.c_double()
        _check(fn(self._h, ctypes.byref(score), _plane_ptrs(ref), _plane_ptrs(dist), stride, stride2), fn.__name__)
        return score.value

    def compute_float(self, ref, dist):
        """ref/dist: three 2-D float32 planes (R, G, B), gamma-encoded in [0, 1]."""
        ref = [np.asarray(p, np.float32) for p in ref]
        dist = [np.asarray(p, np.float32) for p in dist]
        return self._compute(lib().Vship_ComputeSSIMU2Floatv2, ref, dist, ref[0].strides[0], dist[0].strides[0])

    def compute_uint16(self, ref, dist):
        """ref/dist: three 2-D uint16 planes, full range."""
        ref = [np.asarray(p, np.uint16) for p in ref]
        dist = [np.asarray(p, np.uint16) for p in dist]
        return self._compute(lib().Vship_ComputeSSIMU2Uint16v2, ref, dist, ref[0].strides[0], dist[0].strides[0])

    def close(self):
        if self._h is not None:
            _check(lib().Vship_SSIMU2Free(self._h), "Vship_SSIMU2Free")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class DevicePairs:
    """Host frame pairs copied into one device allocation, with the
    Vship_SSIMU2Pair descriptors that point into it."""

    def __init__(self, pairs, fmt, bits=8, matrix=MATRIX_BT709, chroma_location=CHROMA_LEFT):
        """pairs: list of (ref_planes, dist_planes), each three 2-D numpy arrays."""
        offs, total = [], 0
        for ref, dist in pairs:
            o = []
            for p in list(ref) + list(dist):
                o.append(total)
                total += (np.ascontiguousarray(p).nbytes + 255) // 256 * 256
            offs.append(o)
        self.buffer = DeviceBuffer(total)
        self.descs = (Pair * len(pairs))()
        for i, (ref, dist) in enumerate(pairs):
            planes = [np.ascontiguousarray(p) for p in list(ref) + list(dist)]
            for j, p in enumerate(planes):
                self.buffer.upload(p, offs[i][j])
            d = self.descs[i]
            for j in range(3):
                d.ref[j] = self.buffer.ptr + offs[i][j]
                d.dist[j] = self.buffer.ptr + offs[i][3 + j]
                d.ref_stride[j] = planes[j].strides[0]
                d.dist_stride[j] = planes[3 + j].strides[0]
            d.height, d.width = planes[0].shape
            d.format = fmt
            d.bits = bits
            d.matrix = matrix
            d.chroma_location = chroma_location
        self.n = len(pairs)
        self.nbytes = total

    def free(self):
        self.buffer.free()


class HostPairs:
    """Frame pairs in host memory with the Vship_SSIMU2Pair descriptors that
    point at them, for SSIMU2Batch.submit_host. pinned=True packs them into
    one page-locked allocation (Vship_PinnedMalloc), which submit_host
    uploads by DMA directly; pinned=False keeps them in pageable numpy memory,
    which goes through the handler's page-locked staging ring."""

    def __init__(self, pairs, fmt, bits=8, matrix=MATRIX_BT709, chroma_location=CHROMA_LEFT, pinned=False):
        self.n = len(pairs)
        self.descs = (Pair * self.n)()
        self.pinned = None
        self._keep = []
        # pageable planes keep their row pitch (rows must be contiguous);
        # page-locked ones are packed
        planes_all = [[p if p.ndim == 2 and p.strides[1] == p.itemsize and p.strides[0] > 0 else
                       np.ascontiguousarray(p) for p in (np.asarray(q) for q in list(ref) + list(dist))]
                      for ref, dist in pairs]
        if pinned:
            total = sum((p.nbytes + 255) // 256 * 256 for planes in planes_all for p in planes)
            self.pinned = PinnedBuffer(total)
            off = 0
            placed = []
            for planes in planes_all:
                q = []
                for p in planes:
                    v = self.pinned.array[off:off + p.nbytes].view(p.dtype).reshape(p.shape)  # packed rows
                    v[...] = p
                    q.append(v)
                    off += (p.nbytes + 255) // 256 * 256
                placed.append(q)
            planes_all = placed
        else:
            self._keep = planes_all
        self.nbytes = sum(p.nbytes for planes in planes_all for p in planes)
        for i, planes in enumerate(planes_all):
            d = self.descs[i]
            for j in range(3):
                d.ref[j] = planes[j].ctypes.data
                d.dist[j] = planes[3 + j].ctypes.data
                d.ref_stride[j] = planes[j].strides[0]
                d.dist_stride[j] = planes[3 + j].strides[0]
            d.height, d.width = planes[0].shape
            d.format = fmt
            d.bits = bits
            d.matrix = matrix
            d.chroma_location = chroma_location

    def free(self):
        if self.pinned is not None:
            self.pinned.free()
            self.pinned = None
        self._keep = []


class SSIMU2Batch:
    """Batched SSIMULACRA2 over device-resident pairs (Vship_SSIMU2Batch*);
    submit_host takes pairs in host memory. Two batches may be outstanding:
    wait() returns the oldest."""

    def __init__(self, max_pairs, max_width, max_height, fp16=False, arena_bytes=0, map_tile_height=0,
                 persistent_grid=0, use_graph=False, profile=False):
        cfg = BatchConfig(max_pairs=max_pairs, max_width=max_width, max_height=max_height,
                          fp16_planes=int(bool(fp16)), arena_bytes=int(arena_bytes),
                          map_tile_height=int(map_tile_height), persistent_grid=int(persistent_grid),
                          use_graph=int(bool(use_graph)), profile=int(bool(profile)))
        self.max_pairs = max_pairs
        self._ns = []
        self._h = Handler()
        _check(lib().Vship_SSIMU2BatchInit(ctypes.byref(self._h), ctypes.byref(cfg)), "Vship_SSIMU2BatchInit")

    def submit(self, descs, n=None):
        n = len(descs) if n is None else n
        _check(lib().Vship_SSIMU2BatchSubmit(self._h, descs, n), "Vship_SSIMU2BatchSubmit")
        self._ns.append(n)

    def submit_host(self, descs, n=None):
        """descs: Vship_SSIMU2Pair descriptors with host plane pointers (HostPairs.descs)."""
        n = len(descs) if n is None else n
        _check(lib().Vship_SSIMU2BatchSubmitHost(self._h, descs, n), "Vship_SSIMU2BatchSubmitHost")
        self._ns.append(n)

    def host_staging(self, chunk_bytes, slots, copy_threads):
        _check(lib().Vship_SSIMU2BatchHostStaging(self._h, int(chunk_bytes), int(slots), int(copy_threads)),
               "Vship_SSIMU2BatchHostStaging")

    def wait(self, features=False):
        n = self._ns[0] if self._ns else 1
        scores = np.zeros(n)
        feats = np.zeros((n, 108)) if features else None
        _check(lib().Vship_SSIMU2BatchWait(self._h, scores.ctypes.data, feats.ctypes.data if features else None),
               "Vship_SSIMU2BatchWait")
        self._ns.pop(0)
        return (scores, feats) if features else scores

    def compute(self, descs, n=None, features=False):
        """Vship_SSIMU2BatchCompute: refused while a batch is outstanding."""
        n = len(descs) if n is None else n
        scores = np.zeros(n)
        feats = np.zeros((n, 108)) if features else None
        _check(lib().Vship_SSIMU2BatchCompute(self._h, descs, n, scores.ctypes.data,
                                              feats.ctypes.data if features else None), "Vship_SSIMU2BatchCompute")
        return (scores, feats) if features else scores

    def stats(self):
        s = BatchStats()
        _check(lib().Vship_SSIMU2BatchGetStats(self._h, ctypes.byref(s)), "Vship_SSIMU2BatchGetStats")
        return {k: getattr(s, k) for k, _ in BatchStats._fields_}

    def close(self):
        if self._h is not None:
            _check(lib().Vship_SSIMU2BatchFree(self._h), "Vship_SSIMU2BatchFree")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class SSIMU2Server:
    """Continuous-batching server over device-resident clips (Vship_SSIMU2Server*)."""

    def __init__(self, memory_budget, max_pairs_per_step, max_width, max_height, max_inflight_clips=0,
                 fp16=False, map_tile_height=0, persistent_grid=0, use_graph=False):
        cfg = ServerConfig(memory_budget=int(memory_budget), max_pairs_per_step=max_pairs_per_step,
                           max_inflight_clips=max_inflight_clips, max_width=max_width, max_height=max_height,
                           fp16_planes=int(bool(fp16)), map_tile_height=map_tile_height,
                           persistent_grid=persistent_grid, use_graph=int(bool(use_graph)))
        self._h = Handler()
        _check(lib().Vship_SSIMU2ServerInit(ctypes.byref(self._h), ctypes.byref(cfg)), "Vship_SSIMU2ServerInit")
        self._n = {}

    @staticmethod
    def clock_ms():
        return lib().Vship_SSIMU2ServerClockMs()

    def submit(self, descs, n=None, arrival_ms=0.0):
        """descs: ctypes array of Pair (device planes that stay valid until the clip finishes)."""
        n = len(descs) if n is None else n
        cid = ctypes.c_int()
        _check(lib().Vship_SSIMU2ServerSubmitClip(self._h, descs, n, float(arrival_ms), ctypes.byref(cid)),
               "Vship_SSIMU2ServerSubmitClip")
        self._n[cid.value] = n
        return cid.value

    def poll(self, clip_id):
        """(scores, latency_ms) once finished, else None."""
        n = self._n[clip_id]
        scores = np.zeros(n)
        done = ctypes.c_int()
        lat = ctypes.c_double()
        _check(lib().Vship_SSIMU2ServerPollClip(self._h, clip_id, ctypes.byref(done), scores.ctypes.data,
                                                ctypes.byref(lat)), "Vship_SSIMU2ServerPollClip")
        if not done.value:
            return None
        del self._n[clip_id]
        return scores, lat.value

    def wait(self, clip_id):
        n = self._n.pop(clip_id)
        scores = np.zeros(n)
        lat = ctypes.c_double()
        _check(lib().Vship_SSIMU2ServerWaitClip(self._h, clip_id, scores.ctypes.data, ctypes.byref(lat)),
               "Vship_SSIMU2ServerWaitClip")
        return scores, lat.value

    def stats(self):
        st = ServerStats()
        _check(lib().Vship_SSIMU2ServerGetStats(self._h, ctypes.byref(st)), "Vship_SSIMU2ServerGetStats")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self._h is not None:
            _check(lib().Vship_SSIMU2ServerFree(self._h), "Vship_SSIMU2ServerFree")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def autotune(descs, n=None, fp16=False, repeats=3):
    """Times every launch shape on the sample batch (Vship_SSIMU2BatchTune); the
    best one is recorded for the sample's size class. Returns a dict."""
    n = len(descs) if n is None else n
    r = TuneResult()
    _check(lib().Vship_SSIMU2BatchTune(descs, n, int(bool(fp16)), int(repeats), ctypes.byref(r)),
           "Vship_SSIMU2BatchTune")
    table = [{"map_tile_height": c.map_tile_height, "persistent_grid": c.persistent_grid,
              "use_graph": c.use_graph, "ms": c.ms} for c in r.candidates[:r.ncandidates]]
    return {"map_tile_height": r.map_tile_height, "persistent_grid": r.persistent_grid, "use_graph": r.use_graph,
            "best_ms": r.best_ms, "candidates": table}


def tune_save(path):
    _check(lib().Vship_SSIMU2TuneSave(str(path).encode()), "Vship_SSIMU2TuneSave")


def tune_load(path):
    _check(lib().Vship_SSIMU2TuneLoad(str(path).encode()), "Vship_SSIMU2TuneLoad")


class ButteraugliHandler:
    """Per-frame Butteraugli handle of fixed size: the reference's
    Vship_ButteraugliInit[v2] / Vship_ComputeButteraugli{Float,Uint16}[v2] /
    Vship_ButteraugliFree (src/VshipAPI.h:113-148, src/VshipLib.cpp:252-365).
    The score's fields are ints, the norms truncated, as there."""

    def __init__(self, width, height, qnorm=2, intensity_multiplier=203.0):
        self.width, self.height = int(width), int(height)
        self._h = Handler()
        code = lib().Vship_ButteraugliInitv2(ctypes.byref(self._h), self.width, self.height, int(qnorm),
                                             float(intensity_multiplier))
        if code != 0:
            lib().Vship_ButteraugliFree(self._h)
            raise VshipError(code, "Vship_ButteraugliInitv2")

    def _compute(self, fn, ref, dist, distmap):
        score = ButterScore()
        dm = np.zeros((self.height, self.width), np.float32) if distmap else None
        _check(fn(self._h, ctypes.byref(score), dm.ctypes.data if distmap else None,
                  dm.strides[0] if distmap else 0, _plane_ptrs(ref), _plane_ptrs(dist), ref[0].strides[0],
                  dist[0].strides[0]), fn.__name__)
        norms = (score.norm2, score.norm3, score.norminf)
        return (norms, dm) if distmap else norms

    def compute_float(self, ref, dist, distmap=False):
        """ref/dist: three 2-D float32 planes (R, G, B), gamma-encoded in [0, 1]."""
        ref = [np.asarray(p, np.float32) for p in ref]
        dist = [np.asarray(p, np.float32) for p in dist]
        return self._compute(lib().Vship_ComputeButteraugliFloatv2, ref, dist, distmap)

    def compute_uint16(self, ref, dist, distmap=False):
        """ref/dist: three 2-D uint16 planes, full range."""
        ref = [np.asarray(p, np.uint16) for p in ref]
        dist = [np.asarray(p, np.uint16) for p in dist]
        return self._compute(lib().Vship_ComputeButteraugliUint16v2, ref, dist, distmap)

    def close(self):
        if self._h is not None:
            _check(lib().Vship_ButteraugliFree(self._h), "Vship_ButteraugliFree")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def butter_pair_footprint(w, h):
    return int(lib().Vship_ButteraugliPairFootprint(int(w), int(h)))


class ButterDevicePairs:
    """Host frame pairs (RGB planes, u16 or f32) copied into one device
    allocation, with Vship_ButteraugliPair descriptors pointing into it and,
    optionally, a device distortion map per pair."""

    def __init__(self, pairs, distmaps=False, fmt=None, bits=8, matrix=MATRIX_BT709, chroma_location=CHROMA_LEFT):
        """pairs: list of (ref_planes, dist_planes). fmt: a PF_* code; by
        default PF_RGB_U16 for uint16 planes, PF_RGB_F32 otherwise."""
        offs, total = [], 0
        h, w = np.asarray(pairs[0][0][0]).shape
        for ref, dist in pairs:
            o = []
            for p in list(ref) + list(dist):
                o.append(total)
                total += (np.ascontiguousarray(p).nbytes + 255) // 256 * 256
            offs.append(o)
        map_bytes = (w * h * 4 + 255) // 256 * 256
        self.map_offset = total
        if distmaps:
            total += map_bytes * len(pairs)
        self.buffer = DeviceBuffer(total)
        self.descs = (ButterPair * len(pairs))()
        self.width, self.height, self.map_bytes = w, h, map_bytes
        for i, (ref, dist) in enumerate(pairs):
            planes = [np.ascontiguousarray(p) for p in list(ref) + list(dist)]
            for j, p in enumerate(planes):
                self.buffer.upload(p, offs[i][j])
            d = self.descs[i]
            for j in range(3):
                d.ref[j] = self.buffer.ptr + offs[i][j]
                d.dist[j] = self.buffer.ptr + offs[i][3 + j]
                d.ref_stride[j] = planes[j].strides[0]
                d.dist_stride[j] = planes[3 + j].strides[0]
            if distmaps:
                d.distmap = self.buffer.ptr + self.map_offset + i * map_bytes
                d.distmap_stride = w * 4
            d.bits = bits
            d.matrix = matrix
            d.chroma_location = chroma_location
        if fmt is None:
            fmt = PF_RGB_U16 if np.asarray(pairs[0][0][0]).dtype == np.uint16 else PF_RGB_F32
        self.format = fmt
        self.n = len(pairs)
        self.distmaps = distmaps

    def distmap(self, i):
        if not self.distmaps:
            raise ValueError("no distortion maps requested")
        raw = self.buffer.download(self.width * self.height * 4, self.map_offset + i * self.map_bytes)
        return raw.view(np.float32).reshape(self.height, self.width).copy()

    def free(self):
        self.buffer.free()


class ButteraugliBatch:
    """Batched Butteraugli over device-resident pairs (Vship_ButteraugliBatch*)."""

    def __init__(self, width, height, max_pairs=1, qnorm=2, intensity_multiplier=203.0, profile=False):
        cfg = ButterBatchConfig(width=int(width), height=int(height), max_pairs=int(max_pairs), qnorm=int(qnorm),
                                intensity_multiplier=float(intensity_multiplier), profile=int(bool(profile)))
        self._h = Handler()
        code = lib().Vship_ButteraugliBatchInit(ctypes.byref(self._h), ctypes.byref(cfg))
        if code != 0:
            self._h = None
            raise VshipError(code, "Vship_ButteraugliBatchInit")

    def compute(self, dp, n=None):
        """dp: ButterDevicePairs. Returns an (n, 3) array of Q-norm, 3-norm, inf-norm."""
        n = dp.n if n is None else n
        norms = np.zeros((n, 3))
        _check(lib().Vship_ButteraugliBatchCompute(self._h, dp.format, dp.descs, n, norms.ctypes.data),
               "Vship_ButteraugliBatchCompute")
        return norms

    def stats(self):
        s = ButterBatchStats()
        _check(lib().Vship_ButteraugliBatchGetStats(self._h, ctypes.byref(s)), "Vship_ButteraugliBatchGetStats")
        out = {k: getattr(s, k) for k, _ in ButterBatchStats._fields_ if k != "stage_bytes"}
        out["stage_bytes"] = {BUTTER_STAGES[i]: int(s.stage_bytes[i]) for i in range(8)}
        return out

    def close(self):
        if self._h is not None:
            _check(lib().Vship_ButteraugliBatchFree(self._h), "Vship_ButteraugliBatchFree")
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

