"""HIP-event timing over the C ABI (utils::CUDATimer, common/Timer.h:38-64)."""
from __future__ import annotations

import ctypes as C

from . import _lib
from ._lib import check, lib

_P = C.c_void_p
_lib.bind("gary_timer_create", C.c_int, [C.POINTER(_P)])
_lib.bind("gary_timer_destroy", C.c_int, [_P])
_lib.bind("gary_timer_start", C.c_int, [_P, _P])
_lib.bind("gary_timer_stop", C.c_int, [_P, _P, C.POINTER(C.c_float)])


class EventTimer:
    """start()/stop() record HIP events on `stream`; stop() returns elapsed milliseconds."""

    def __init__(self, stream: int | None = None):
        self.stream = stream
        h = _P(None)
        check(lib.gary_timer_create(C.byref(h)))
        self._h = h

    def start(self) -> None:
        check(lib.gary_timer_start(self._h, self.stream))

    def stop(self) -> float:
        ms = C.c_float(0)
        check(lib.gary_timer_stop(self._h, self.stream, C.byref(ms)))
        return float(ms.value)

    def __del__(self):
        try:
            if self._h.value:
                lib.gary_timer_destroy(self._h)
        except Exception:
            pass


_lib.bind("gary_kernel_timing", C.c_int, [C.c_int])
_lib.bind("gary_kernel_timing_read", C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_float)])


def kernel_timing(enable: bool) -> None:
    """Turn per-launch event timing of the dominant counting kernel on/off (clears the record)."""
    check(lib.gary_kernel_timing(1 if enable else 0))


def kernel_timing_read() -> tuple[int, float]:
    """(launches, summed ms) of the dominant counting kernel since the last read or enable."""
    n, ms = C.c_int(0), C.c_float(0)
    check(lib.gary_kernel_timing_read(C.byref(n), C.byref(ms)))
    return n.value, float(ms.value)
