"""vship_amd — MI355X (gfx950) SSIMULACRA2 behind Vship's C ABI.

The product is the native library ``vship_amd/libvship.so`` (hand-written HIP
kernels + C++ host engine). This package binds it with ctypes and mirrors the
reference's API surface for this path:

* :class:`SSIMU2Handler` — ``Vship_SSIMU2Init`` / ``Vship_ComputeSSIMU2*`` /
  ``Vship_SSIMU2Free`` (the reference's src/VshipAPI.h:87-111);
* :class:`SSIMU2Batch` — the additive batched entry points
  (``include/VshipSSIMU2Batch.h``) over device-resident frame pairs;
* :class:`SSIMU2Server` — the continuous-batching request server;
* :class:`DeviceBuffer` — device memory through the library's helpers.

Nothing here falls back to a CPU implementation: every call goes through the
HIP library, and a missing library raises :class:`VshipNotBuilt`.
"""
from .api import (  # noqa: F401
    ERRORS,
    DeviceBuffer,
    PinnedBuffer,
    SSIMU2Batch,
    SSIMU2Handler,
    SSIMU2Server,
    VshipError,
    VshipNotBuilt,
    device_count,
    device_info,
    error_message,
    gpu_full_check,
    lib,
    lib_path,
    pair_footprint,
    set_device,
    version,
)
