"""ctypes binding of libgary.so (the C ABI declared in include/gary.h).

The library is the product path: if it is missing, importing this module raises. There is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libgary.so"

GARY_OK = 0
GARY_ERR_INVALID = -1
GARY_ERR_HIP = -2
GARY_ERR_OOM = -3
GARY_ERR_OVERFLOW = -4
GARY_ERR_UNSUPPORTED = -5
GARY_ERR_STATE = -6

CSR_ORIENT_DEGREE = 0
CSR_SYMMETRIC = 1
CSR_ORIENT_ID = 2


class GaryError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class DeviceCSR(C.Structure):
    _fields_ = [("m", C.c_int32), ("nnz", C.c_int64), ("indPtr", C.c_void_p),
                ("colIdx", C.c_void_p), ("deg", C.c_void_p)]


class DeviceCOO(C.Structure):
    _fields_ = [("m", C.c_int32), ("nnz", C.c_int64), ("rowIdx", C.c_void_p),
                ("colIdx", C.c_void_p), ("deg", C.c_void_p)]


class DeviceGraph(C.Structure):
    _fields_ = [("csr", DeviceCSR), ("coo", DeviceCOO)]


class LaunchCfg(C.Structure):
    _fields_ = [("block", C.c_int32), ("grid", C.c_int32), ("group", C.c_int32),
                ("cache_depth", C.c_int32), ("partition", C.c_int32), ("bin_small", C.c_int32),
                ("bin_hub", C.c_int32), ("direction", C.c_int32)]


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built: run `python -m gary_amd.build` "
                          "(or __graft_entry__.build()) first")
    return C.CDLL(str(LIB_PATH))


lib = _load()

_P = C.c_void_p
_I32, _I64, _U64, _D = C.c_int32, C.c_int64, C.c_uint64, C.c_double
_SIGS = {
    "gary_last_error": (C.c_char_p, []),
    "gary_version": (C.c_int, []),
    "gary_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "gary_device_synchronize": (C.c_int, [C.c_int]),
    "gary_gen_er": (C.c_int, [_I32, _I64, _U64, _P, _P]),
    "gary_gen_rmat": (C.c_int, [_I32, _I64, _D, _D, _D, _U64, _P, _P]),
    "gary_gen_chung_lu": (C.c_int, [_I32, _I64, _D, _D, _U64, _P, _P]),
    "gary_csr_from_edges": (C.c_int, [_I32, _I64, _P, _P, C.c_int, _P, _P, _I64, _P]),
    "gary_malloc": (C.c_int, [C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]),
    "gary_free": (C.c_int, [C.c_int, _P]),
    "gary_memcpy_h2d": (C.c_int, [C.c_int, _P, _P, C.c_size_t]),
    "gary_memcpy_d2h": (C.c_int, [C.c_int, _P, _P, C.c_size_t]),
    "gary_memcpy_d2d": (C.c_int, [C.c_int, _P, _P, C.c_size_t]),
    "gary_memset": (C.c_int, [C.c_int, _P, C.c_int, C.c_size_t]),
    "gary_stream_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "gary_stream_destroy": (C.c_int, [C.c_int, _P]),
    "gary_stream_synchronize": (C.c_int, [C.c_int, _P]),
}


def _bind(name, restype, argtypes):
    fn = getattr(lib, name, None)
    if fn is None:
        return None
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


for _n, (_r, _a) in _SIGS.items():
    _bind(_n, _r, _a)


def bind(name: str, restype, argtypes):
    """Declare the signature of an exported symbol and return it."""
    fn = _bind(name, restype, argtypes)
    if fn is None:
        raise AttributeError(f"libgary.so does not export {name}")
    return fn


def last_error() -> str:
    return (lib.gary_last_error() or b"").decode(errors="replace")


def check(code: int) -> int:
    if code != GARY_OK:
        raise GaryError(code, last_error())
    return code


def ptr(a: np.ndarray | None) -> int | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data


def device_count() -> int:
    c = C.c_int(0)
    check(lib.gary_device_count(C.byref(c)))
    return c.value


def exported_symbols() -> list[str]:
    """Names of the gary_* functions declared in include/gary.h."""
    import re
    hdr = Path(__file__).resolve().parent.parent / "include" / "gary.h"
    text = hdr.read_text()
    return sorted(set(re.findall(r"\b(gary_[a-z0-9_]+)\s*\(", text)))

