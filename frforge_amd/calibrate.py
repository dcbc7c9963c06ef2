"""A plain device-to-device copy of known size through the same HIP runtime as the library, for
calibrating rocprofv3's FETCH_SIZE / WRITE_SIZE counters against known HBM bytes."""
from __future__ import annotations

import ctypes as C

from .device import load_library

_HIP = "libamdhip64.so.7"


def device_copy(nbytes: int, repeats: int = 3) -> int:
    load_library()                       # the library's runtime is the one already loaded
    hip = C.CDLL(_HIP)
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    try:
        hip.hipMemset(src, 1, C.c_size_t(nbytes))
        for _ in range(repeats):
            hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
        hip.hipDeviceSynchronize()
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    return nbytes
