"""The C-ABI library loads, exports every symbol include/srk.h declares, and
rejects bad arguments before touching the GPU (CPU only)."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from shrinkage_full_tests_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "srk.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(\w+)\s*\(", src, flags=re.M)))


def test_header_declares_the_reference_entry_points():
    names = _declared()
    for n in ("srkTrainingForward", "adjointSRK", "shrinkageTangentOaS",
              "shrinkageTangentAdjointOaS", "srkForwardOaS", "srkPrecomputeKernelFreq"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with srk.h"


def test_error_strings():
    L = _lib.lib()
    for code in range(0, -10, -1):
        s = L.srkGetErrorString(code).decode()
        assert s and s != "unknown status"
    assert L.srkGetErrorString(-99).decode() == "unknown status"


def _create(N, W, e, M, B=1):
    h = ctypes.c_void_p()
    st = _lib.lib().srkWorkspaceCreate(N, W, e, M, B, 0, ctypes.byref(h))
    return st, h


def test_workspace_argument_errors():
    assert _create(256, 200, 15, 256)[0] == -2        # W != N - 2e
    assert _create(256, 256, 0, 256)[0] == -2         # erosion < 1
    assert _create(30, -2, 16, 256)[0] == -2          # W <= 0
    assert _create(256, 226, 15, 300)[0] == -5        # M not supported
    assert _create(1024, 904, 60, 128)[0] == -5       # M not supported
    assert _create(1024, 770, 127, 256)[0] == -5      # M <= 2e + 2
    assert _create(256, 226, 15, 256, B=0)[0] == -7   # B_max < 1
    assert _lib.lib().srkWorkspaceCreate(256, 226, 15, 256, 1, 0, None) == -3


def test_entry_points_reject_null_workspace():
    L = _lib.lib()
    assert L.srkTrainingForward(1, 1, 8, 8, 256, 226, 15, 1.0, 1, None, 1, None) == -3
    assert L.srkForwardBatched(1, 1, 8, 8, 256, 226, 15, 1.0, 1, None, 1, None) == -3
    assert L.srkAdjointBatched(1, 1, 1, 8, 8, 256, 226, 15, 1.0, 1, None, 1, 1, None) == -3
    assert L.shrinkageTangentOaS(1, 1, 1, 1, 8, 8, 256, 226, 15, 1.0, None, 1, None) == -3
    assert L.srkForwardBatchedU(1, 1, 8, 8, 256, 226, 15, 1.0, 1, None, 1, 1, 1, None) == -3
    assert L.srkAdjointBatchedU(1, 1, 1, 1, 8, 8, 256, 226, 15, 1.0, 1, None, 1, 1, None) == -3


def test_kernel_spectra_argument_errors():
    L = _lib.lib()
    assert L.srkKernelSpectraSRSRK(30, 5.0, 1.0, 3.0, 256, None, None, None, None, None) == -9
    assert L.srkKernelSpectraSRSRK(31, -1.0, 1.0, 3.0, 256, None, None, None, None, None) == -9
    # any spectrum size (the reference plans cuFFT at any N), but not below the kernel width
    assert L.srkKernelSpectraSRSRK(31, 5.0, 1.0, 3.0, 16, None, None, None, None, None) == -5
    assert L.srkKernelSpectraSRSRK(31, 5.0, 1.0, 3.0, 300, None, None, None, None, None) == 0
    assert L.srkPrecomputeKernelFreq(None, 31, 31, 256, None, None) == -3


@pytest.mark.parametrize("kw,sigma", [(31, 5.0), (61, 10.0), (121, 20.0)])
def test_product_kernel_generation_matches_scipy(kw, sigma):
    """The library's own FP64 K1 / Green's tensor (host code) against scipy goldens."""
    g = json.load(open(os.path.join(ROOT, "tests", "golden", "k1_scipy.json")))
    hx = np.zeros((kw, kw), np.float32)
    hy = np.zeros((kw, kw), np.float32)
    st = _lib.lib().srkKernelSpectraSRSRK(kw, sigma, 1.0, 3.0, 256, None, None,
                                          hx.ctypes.data, hy.ctypes.data, None)
    assert st == 0
    for i, j, kx, ky in g["srsrk_samples"][f"{kw}_{sigma}"]:
        assert abs(hx[i, j] - kx) <= 1e-6 * abs(kx) + 1e-30
        assert abs(hy[i, j] - ky) <= 1e-6 * abs(ky) + 1e-30


def test_kernel_kind_names_and_timing_null_workspace():
    L = _lib.lib()
    names = [L.srkKernelKindName(k).decode() for k in range(7)]
    assert names == ["k_p1_cols", "k_p1_adj_src", "k_p2_conv", "k_p2_corr", "k_p3_fwd", "k_p3_adj",
                     "k_fused"]
    assert L.srkKernelKindName(7) == b"" and L.srkKernelKindName(-1) == b""
    assert L.srkWorkspaceSetTiming(None, 1) == -3
    assert L.srkWorkspaceKernelTimes(None, None, None, None) == -3


def test_cd_entry_points_reject_bad_arguments():
    L = _lib.lib()
    # null pointers, bad batch, bad sizes: rejected before any GPU work
    assert L.srkMeasureCDBatched(None, 1, 64, 16, 8, 0.0, None, None, None, None) == -3
    p = 256  # any non-null value: the checks below fail first
    assert L.srkMeasureCDBatched(p, 0, 64, 16, 8, 0.0, p, p, p, None) == -7
    assert L.srkMeasureCDBatched(p, 1, 1, 16, 8, 0.0, p, p, p, None) == -2
    assert L.srkMeasureCDBatched(p, 1, 50000, 16, 8, 0.0, p, p, p, None) == -5
    assert L.srkShrinkageCDBatched(p, p, 256, 100, 60, 1, 1.0, 1, 16, 8, 0.0, p, p, p, None) == -2


def test_host_tensors_rejected_before_the_library():
    import torch

    import shrinkage_full_tests_amd as srk

    class _Ws:           # stands in for a workspace; no device is touched
        M, handle = 256, None

    E = torch.zeros((1, 256, 256))
    with pytest.raises(ValueError, match="CUDA"):
        srk.forward_batched(E, E, None, None, 15, 1.0, _Ws())
    with pytest.raises(ValueError, match="CUDA"):
        srk.shrinkageTangentOaS(E[0], E[0], E[0], E[0], None, None, 256, 226, 15, 1.0, _Ws(),
                                torch.zeros((226, 226)))
