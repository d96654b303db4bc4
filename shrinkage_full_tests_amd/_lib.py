"""ctypes binding of libsrk_amd.so (the C ABI declared in include/srk.h)."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsrk_amd.so")

_lib = None

_i = ctypes.c_int
_f = ctypes.c_float
_d = ctypes.c_double
_p = ctypes.c_void_p
_sz = ctypes.c_size_t

# name -> (restype, argtypes); pointers travel as integers (device addresses)
SIGNATURES = {
    "srkGetErrorString": (ctypes.c_char_p, [_i]),
    "srkLastHipError": (_i, []),
    "srkGetVersion": (_i, []),
    "srkSynchronize": (_i, [_p]),
    "srkPrecomputeKernelFreq": (_i, [_p, _i, _i, _i, _p, _p]),
    "srkKernelSpectraSRSRK": (_i, [_i, _d, _d, _d, _i, _p, _p, _p, _p, _p]),
    "srkKernelSpectraSRSRKDevice": (_i, [_i, _d, _d, _d, _i, _p, _p, _p, _p, _p]),
    "srkWorkspaceCreate": (_i, [_i, _i, _i, _i, _i, _sz, ctypes.POINTER(_p)]),
    "srkWorkspaceDestroy": (_i, [_p]),
    "srkWorkspaceBytes": (_sz, [_p]),
    "srkWorkspaceSetChunk": (_i, [_p, _i]),
    "srkWorkspaceSetPrecision": (_i, [_p, _i]),
    "srkWorkspaceSetTuning": (_i, [_p, _i, _i]),
    "srkWorkspaceInfo": (_i, [_p] + [ctypes.POINTER(_i)] * 5),
    "srkWorkspaceSetTiming": (_i, [_p, _i]),
    "srkWorkspaceFusedStatus": (_i, [_p, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "srkWorkspaceKernelTimes": (_i, [_p, _p, _p, _p]),
    "srkKernelKindName": (ctypes.c_char_p, [_i]),
    "srkTrainingForward": (_i, [_p, _p, _p, _p, _i, _i, _i, _f, _i, _p, _p, _p]),
    "srkCheckSpectraSupport": (_i, [_p, _p, _i, _i, _p, _p, ctypes.POINTER(_d)]),
    "srkForwardOaS": (_i, [_p, _p, _p, _p, _i, _i, _i, _f, _p, _p, _p, _p, _p]),
    "srkForwardOaSNoAlloc": (_i, [_p, _p, _p, _p, _i, _i, _i, _f, _i, _p, _p, _p, _p]),
    "adjointSRK": (_i, [_p, _p, _p, _p, _p, _p, _i, _i, _i, _f, _p, _p, _p, _p]),
    "shrinkageTangentOaS": (_i, [_p, _p, _p, _p, _p, _p, _i, _i, _i, _f, _p, _p, _p]),
    "shrinkageTangentAdjointOaS": (_i, [_p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _f, _p,
                                        _p, _p, _p]),
    "srkForwardBatched": (_i, [_p, _p, _p, _p, _i, _i, _i, _f, _i, _p, _p, _p]),
    "srkAdjointBatched": (_i, [_p, _p, _p, _p, _p, _i, _i, _i, _f, _i, _p, _p, _p, _p]),
    "srkForwardBatchedU": (_i, [_p, _p, _p, _p, _i, _i, _i, _f, _i, _p, _p, _p, _p, _p]),
    "srkAdjointBatchedU": (_i, [_p, _p, _p, _p, _p, _p, _i, _i, _i, _f, _i, _p, _p, _p, _p]),
    "srkTangentBatched": (_i, [_p, _p, _p, _p, _p, _p, _i, _i, _i, _f, _i, _p, _p, _p]),
    "srkTangentAdjointBatched": (_i, [_p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _f, _i, _p,
                                      _p, _p, _p]),
    "srkMeasureCDBatched": (_i, [_p, _i, _i, _i, _i, _f, _p, _p, _p, _p]),
    "srkMeasureCDBatchedEx": (_i, [_p, _p, _p, _i, _i, _i, _i, _p, _p, _p, _p]),
    "srkBuildPolyphaseTable": (_i, [_i, _i, _p, _p]),
    "srkUpsampleBatched": (_i, [_p, _p, _i, _i, _i, _i, _p, _p]),
    "srkGpGenerate": (_i, [_d, _d, _d, _d, _i, _i, ctypes.c_ulonglong, _i, _p]),
    "srkShrinkageCDBatched": (_i, [_p, _p, _i, _i, _i, _i, _f, _i, _i, _i, _f, _p, _p, _p, _p]),
}


class SrkError(RuntimeError):
    """Raised when a C-ABI call returns a negative status."""

    def __init__(self, status: int, where: str):
        msg = lib().srkGetErrorString(status).decode()
        if status == -8:
            msg += f" (hipError {lib().srkLastHipError()})"
        super().__init__(f"{where}: {msg} [status {status}]")
        self.status = status


def lib():
    """Load libsrk_amd.so; raise if it was not built (there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m shrinkage_full_tests_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, where: str) -> None:
    if status != 0:
        raise SrkError(status, where)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
