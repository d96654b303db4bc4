"""Differentiated single-image path at N in {1024, 2048, 4096}, KW = 61 (GPU).

1. Oracle windows: shrinkageTangentOaS and shrinkageTangentAdjointOaS
   (the drop-in entry points, tan_tanadj/srk_tangent.cuh:106-151 and
   tan_tanadj/srk_tangent_adjoint.cuh:100-159) against the FP64 oracle on
   windows placed across overlap-and-save tile seams, at the fp32 tolerance
   (tol_rel 1e-3, tol_abs 1e-5, README.md:8-10).
2. The reference's derivative-check matrix (tan_tanadj/main_tan_tanadj.cu:
   456-648) on every output pixel: tangent modes A (df only), B (dg only) and
   C (both); tangent-adjoint of lambda_f in modes D (dlam_h only), E (dg only)
   and G (all), of lambda_g in modes D, F (df only) and G; three directions
   each, central differences with eps in {1e-2, 1e-1}, pass if some eps
   passes every pixel (tangent_tester.h:454-560).  f, g are Gaussian random
   fields; directions are i.i.d. N(0, 1) with seeds 1, 2, 3.  The operator is
   bilinear, so the central differences carry rounding error only.
"""
import numpy as np
import pytest

import shrinkage_full_tests_amd as srk
from shrinkage_full_tests_amd import patterns
from shrinkage_full_tests_amd import testing as T

pytestmark = pytest.mark.gpu

TOL_REL, TOL_ABS = 1e-3, 1e-5
KW = 61
E = (KW - 1) // 2
SIGMA = E / 3.0


def _torch():
    import torch
    return torch


def assert_close(a, n, what=""):
    err = np.abs(a - n)
    lim = np.maximum(TOL_ABS, TOL_REL * np.maximum(np.abs(a), np.abs(n)))
    bad = err > lim
    assert not bad.any(), (f"{what}: {bad.sum()} violations, max err {err.max():.3e}, "
                           f"max|ref| {np.abs(n).max():.3e}")


def crop(t, X0, Y0, S):
    """[S, S] float64 copy of t[X0:X0+S, Y0:Y0+S] (t: [n, n] tensor), zero outside."""
    n = t.shape[0]
    out = np.zeros((S, S))
    a0, a1 = max(X0, 0), min(X0 + S, n)
    b0, b1 = max(Y0, 0), min(Y0 + S, n)
    if a0 < a1 and b0 < b1:
        out[a0 - X0:a1 - X0, b0 - Y0:b1 - Y0] = t[a0:a1, b0:b1].double().cpu().numpy()
    return out


def _inputs(N):
    torch = _torch()
    W = N - 2 * E
    f = patterns.grf_torch(1, N, seed=7000 + N)[0]
    g = patterns.grf_torch(1, N, seed=8000 + N)[0]
    lam = torch.randn((W, W), generator=torch.Generator(device="cuda").manual_seed(123),
                      device="cuda")
    dirs = []
    for d in (1, 2, 3):
        gen = torch.Generator(device="cuda")
        gen.manual_seed(d)
        dirs.append((torch.randn((N, N), generator=gen, device="cuda"),
                     torch.randn((N, N), generator=gen, device="cuda"),
                     torch.randn((W, W), generator=gen, device="cuda")))
    return f, g, lam, dirs


@pytest.mark.parametrize("N,M", [(1024, 512), (1024, 1024), (2048, 1024), (4096, 1024),
                                 (4096, 2048)])
def test_tangent_pair_windows(orc, N, M):
    torch = _torch()
    W = N - 2 * E
    gx, gy, hx, hy = srk.kernel_spectra_srsrk(KW, SIGMA, M, with_spatial=True)
    Kx, Ky = hx.numpy().astype(np.float64), hy.numpy().astype(np.float64)
    f, g, lam, dirs = _inputs(N)
    df, dg, dlam = dirs[0]
    ws = srk.Workspace(N, W, E, M)
    dh = torch.empty((W, W), device="cuda")
    srk.shrinkageTangentOaS(f, g, df, dg, gx, gy, N, W, E, 1.0, ws, dh)
    a = torch.empty((N, N), device="cuda")
    b = torch.empty((N, N), device="cuda")
    srk.shrinkageTangentAdjointOaS(f, g, df, dg, lam, dlam, gx, gy, N, W, E, 1.0, ws, a, b)
    torch.cuda.synchronize()
    w = 48
    S = M - 2 * E                      # forward tile stride
    wins = [(0, 0), (W // 2, W // 3), (W - w, W - w)]
    if N > M:
        wins.append((S - w // 2, S - w // 2))
    for (ox0, oy0) in wins:
        Sw = w + 2 * E + 2
        cs = [crop(t, ox0, oy0, Sw) for t in (f, g, df, dg)]
        ref = orc.tangent_window(*cs, E, Kx, Ky, 0, 0, w, w)
        got = dh[ox0:ox0 + w, oy0:oy0 + w].double().cpu().numpy()
        assert_close(got, ref, f"dh window at ({ox0},{oy0}) N={N} M={M}")
    Sa = M - 2 * E - 2                 # adjoint tile stride
    qwins = [(0, 0), (N // 2, N // 5), (N - w, N - w)]
    if N > M:
        qwins.append((Sa - w // 2, 2 * Sa - w // 2) if 2 * Sa < N - w else (Sa - w // 2, 0))
    pad = 2 * E + 2
    for (qx0, qy0) in qwins:
        X0, Y0, Sw = qx0 - pad, qy0 - pad, w + 2 * pad
        cs = [crop(t, X0, Y0, Sw) for t in (f, g, df, dg)]
        ls = [crop(t, X0, Y0, Sw - 2 * E) for t in (lam, dlam)]
        ra, rb = orc.tangent_adjoint_window(*cs, *ls, E, Kx, Ky, pad, pad, w, w)
        assert_close(a[qx0:qx0 + w, qy0:qy0 + w].double().cpu().numpy(), ra,
                     f"dLam_f window at ({qx0},{qy0}) N={N} M={M}")
        assert_close(b[qx0:qx0 + w, qy0:qy0 + w].double().cpu().numpy(), rb,
                     f"dLam_g window at ({qx0},{qy0}) N={N} M={M}")


@pytest.mark.parametrize("N,M", [(1024, 512), (2048, 1024), (4096, 1024)])
def test_derivative_check_matrix(N, M):
    """Tangent A/B/C and tangent-adjoint D/E/G (lambda_f), D/F/G (lambda_g) in
    three directions, all output pixels, eps in {1e-2, 1e-1}."""
    torch = _torch()
    W = N - 2 * E
    gx, gy = srk.kernel_spectra_srsrk(KW, SIGMA, M)
    f, g, lam, dirs = _inputs(N)
    ws = srk.Workspace(N, W, E, M)
    cfg = T.TangentConfig(num_samples=-1, epsilons=[1e-2, 1e-1], tol_abs=TOL_ABS, tol_rel=TOL_REL)
    zN = torch.zeros((N, N), device="cuda")
    zW = torch.zeros((W, W), device="cuda")

    def fwd(f_, g_):
        R = torch.empty((W, W), device="cuda")
        return srk.srkForwardOaS(g_, f_, gx, gy, N, W, E, 1.0, ws, R)

    def tan(s, d):
        out = torch.empty((W, W), device="cuda")
        srk.shrinkageTangentOaS(s[0], s[1], d[0] if d[0] is not None else zN,
                                d[1] if d[1] is not None else zN, gx, gy, N, W, E, 1.0, ws, out)
        return out

    def adj(which):
        def fn(f_, g_, l_):
            return srk.adjoint_batched(l_[None], g_[None], f_[None], gx, gy, E, 1.0, ws)[which][0]
        return fn

    def tadj(which):
        def fn(s, d):
            a = torch.empty((N, N), device="cuda")
            b = torch.empty((N, N), device="cuda")
            srk.shrinkageTangentAdjointOaS(
                s[0], s[1], d[0] if d[0] is not None else zN, d[1] if d[1] is not None else zN,
                s[2], d[2] if d[2] is not None else zW, gx, gy, N, W, E, 1.0, ws, a, b)
            return (a, b)[which]
        return fn

    failures = []
    for k, (df, dg, dlam) in enumerate(dirs):
        tan_modes = {"A df-only": (df, None), "B dg-only": (None, dg), "C full": (df, dg)}
        for name, d in tan_modes.items():
            r = T.tangent_test(fwd, tan, [f, g], list(d), cfg)
            if not r.passed:
                failures.append((f"tangent {name} dir={k}", r.best_num_violations,
                                 r.best_abs_error_max))
        adj_modes = {
            (0, "D dlam_h-only"): (None, None, dlam), (0, "E dg-only"): (None, dg, None),
            (0, "G full"): (df, dg, dlam),
            (1, "D dlam_h-only"): (None, None, dlam), (1, "F df-only"): (df, None, None),
            (1, "G full"): (df, dg, dlam)}
        for (which, name), d in adj_modes.items():
            r = T.tangent_adjoint_test(adj(which), tadj(which), [f, g, lam], list(d), cfg)
            if not r.passed:
                failures.append((f"tan-adj {'dLam_f' if which == 0 else 'dLam_g'} {name} dir={k}",
                                 r.best_num_violations, r.best_abs_error_max))
    assert not failures, failures
