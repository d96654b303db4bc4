"""srkForwardOaSNoAlloc (no workspace, static device scratch) against
srkForwardOaS (workspace) through the C ABI: two calls in a row, the second
on a smaller input, at the fp32 tolerance; sizes beyond the scratch fail."""
import ctypes

import numpy as np
import pytest

import shrinkage_full_tests_amd as srk
from shrinkage_full_tests_amd import _lib, patterns

pytestmark = pytest.mark.gpu

TOL_REL, TOL_ABS = 1e-3, 1e-5


def _noalloc(E, I, gx, gy, N, W, e, M, R, ux=None, uy=None, dx=1.0):
    import torch
    L = _lib.lib()
    st = L.srkForwardOaSNoAlloc(E.data_ptr(), I.data_ptr(), gx.data_ptr(), gy.data_ptr(), N, W,
                                e, ctypes.c_float(dx), M,
                                torch.cuda.current_stream().cuda_stream, R.data_ptr(),
                                ux.data_ptr() if ux is not None else None,
                                uy.data_ptr() if uy is not None else None)
    return st


def _close(a, b, what):
    a = a.double().cpu().numpy()
    b = b.double().cpu().numpy()
    err = np.abs(a - b)
    lim = np.maximum(TOL_ABS, TOL_REL * np.maximum(np.abs(a), np.abs(b)))
    assert not (err > lim).any(), (what, float(err.max()))


def test_no_alloc_forward_matches_workspace_forward():
    import torch
    cases = [(4096, 61, 512, 4.0), (300, 31, 256, 1.0)]   # second call: smaller input
    for N, kw, M, dx in cases:
        e = (kw - 1) // 2
        W = N - 2 * e
        gx, gy = srk.kernel_spectra_srsrk(kw, e / 3.0, M)
        E = patterns.grf_torch(1, N, seed=50 + N)[0]
        I = patterns.grf_torch(1, N, seed=60 + N)[0]
        ws = srk.Workspace(N, W, e, M)
        R0 = torch.empty((W, W), device="cuda")
        ux0, uy0 = torch.empty_like(R0), torch.empty_like(R0)
        srk.srkForwardOaS(E, I, gx, gy, N, W, e, dx, ws, R0, ux0, uy0)
        R1 = torch.full((W, W), float("nan"), device="cuda")
        ux1, uy1 = torch.full_like(R1, float("nan")), torch.full_like(R1, float("nan"))
        assert _noalloc(E, I, gx, gy, N, W, e, M, R1, ux1, uy1, dx) == 0
        R2 = torch.full_like(R1, float("nan"))
        assert _noalloc(E, I, gx, gy, N, W, e, M, R2, dx=dx) == 0
        torch.cuda.synchronize()
        for got, want, what in ((R1, R0, "h"), (ux1, ux0, "u_x"), (uy1, uy0, "u_y"),
                                (R2, R0, "h without u")):
            assert bool(torch.isfinite(got).all()), (N, what)
            _close(got, want, f"N={N} M={M} {what}")
        print(f"N={N} M={M}: bit-equal to srkForwardOaS: {bool(torch.equal(R1, R0))}")


def test_no_alloc_rejects_sizes_beyond_scratch():
    import torch
    L = _lib.lib()
    N, e, M = 4096, 30, 2048
    W = N - 2 * e
    t = torch.zeros(8, device="cuda")
    p = t.data_ptr()
    s = torch.cuda.current_stream().cuda_stream
    assert L.srkForwardOaSNoAlloc(p, p, p, p, N, W, e, ctypes.c_float(1.0), M, s, p, None,
                                  None) == -5
    assert L.srkForwardOaSNoAlloc(p, p, p, p, N, W + 1, e, ctypes.c_float(1.0), 512, s, p, None,
                                  None) == -2
