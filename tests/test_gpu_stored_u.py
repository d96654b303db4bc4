"""The batched stored-activation pair (srkForwardBatchedU, srkAdjointBatchedU:
the reference's srkForwardOaS-with-u + adjointSRK flow, batched) against the
FP64 oracle and against the recomputing adjoint."""
import numpy as np
import pytest

import shrinkage_full_tests_amd as srk
from shrinkage_full_tests_amd import patterns

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def assert_close(a, n, tol_rel, tol_abs, what=""):
    err = np.abs(a - n)
    lim = np.maximum(tol_abs, tol_rel * np.maximum(np.abs(a), np.abs(n)))
    bad = err > lim
    assert not bad.any(), (f"{what}: {bad.sum()} violations, max err {err.max():.3e}, "
                           f"max|ref| {np.abs(n).max():.3e}")


@pytest.mark.parametrize("prec", ["fp32", "bf16_exchange"])
@pytest.mark.parametrize("N,kw,M,same", [(512, 121, 512, True), (512, 61, 512, False),
                                         (1024, 61, 512, False)])
def test_stored_u_pair(orc, N, kw, M, same, prec):
    tol = (1e-3, 1e-5) if prec == "fp32" else (1e-2, 1e-3)
    e = (kw - 1) // 2
    W = N - 2 * e
    B = 2
    gx, gy, hx, hy = srk.kernel_spectra_srsrk(kw, e / 3.0, M, with_spatial=True)
    Kx, Ky = hx.numpy().astype(np.float64), hy.numpy().astype(np.float64)
    g = patterns.grf(N, 50 + N, B=B)
    f = g if same else patterns.grf(N, 60 + N, B=B)
    v = np.random.default_rng(kw).standard_normal((B, W, W)).astype(np.float32)
    ws = srk.Workspace(N, W, e, M, B_max=B, precision=prec)
    dg, df, dv = dev(g), dev(f), dev(v)
    R, ux, uy = srk.forward_batched_u(dg, df, gx, gy, e, 1.0, ws)
    R0 = srk.forward_batched(dg, df, gx, gy, e, 1.0, ws)
    lf, lg = srk.adjoint_batched_u(dv, ux, uy, df, gx, gy, e, 1.0, ws)
    lf0, lg0 = srk.adjoint_batched(dv, dg, df, gx, gy, e, 1.0, ws)
    Rh, uxh, uyh, lfh, lgh = host(R), host(ux), host(uy), host(lf), host(lg)
    if prec == "fp32":
        # the same passes write h whether or not u is stored
        assert np.array_equal(Rh, host(R0))
    for b in range(B):
        rux, ruy = orc.conv_valid(g[b], e, Kx, Ky)
        assert_close(uxh[b], rux, 1e-3, 1e-5, f"u_x b{b}")   # u is FP32 in every form
        assert_close(uyh[b], ruy, 1e-3, 1e-5, f"u_y b{b}")
        assert_close(Rh[b], orc.forward(f[b], g[b], e, Kx, Ky, fft=True), *tol, f"h b{b}")
        rf, rg = orc.adjoint(v[b], f[b], g[b], e, Kx, Ky, fft=True)
        assert_close(lfh[b], rf, 1e-3, 1e-5, f"lambda_f b{b}")
        assert_close(lgh[b], rg, *tol, f"lambda_g b{b}")
    assert_close(lfh, host(lf0), *tol, "lambda_f vs recompute")
    assert_close(lgh, host(lg0), *tol, "lambda_g vs recompute")
