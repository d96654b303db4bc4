"""GPU parity of the HIP path (through the C ABI) against the FP64 oracle.

Tolerance (fp32 form, README.md:8-10, stTestCore/adjoint_tester.h:31-38):
per element |a - n| <= max(tol_abs, tol_rel * max(|a|, |n|)),
tol_rel = 1e-3, tol_abs = 1e-5.
"""
import numpy as np
import pytest

import shrinkage_full_tests_amd as srk
from shrinkage_full_tests_amd import patterns

pytestmark = pytest.mark.gpu

TOL_REL, TOL_ABS = 1e-3, 1e-5


def _torch():
    import torch
    return torch


def dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host(t):
    _torch().cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def assert_close(a, n, tol_rel=TOL_REL, tol_abs=TOL_ABS, what=""):
    a = np.asarray(a, np.float64)
    n = np.asarray(n, np.float64)
    assert a.shape == n.shape, (a.shape, n.shape)
    err = np.abs(a - n)
    lim = np.maximum(tol_abs, tol_rel * np.maximum(np.abs(a), np.abs(n)))
    bad = err > lim
    assert not bad.any(), (f"{what}: {bad.sum()} of {bad.size} violations, max err "
                           f"{err.max():.3e}, max|ref| {np.abs(n).max():.3e}")


def kernels(kw, sigma, M):
    gx, gy, hx, hy = srk.kernel_spectra_srsrk(kw, sigma, M, with_spatial=True)
    return gx, gy, hx.numpy().astype(np.float64), hy.numpy().astype(np.float64)


# ---- spectra -----------------------------------------------------------------

@pytest.mark.parametrize("kw,sigma,M", [(31, 5.0, 256), (121, 20.0, 512)])
def test_kernel_spectra_match_numpy_fft(orc, kw, sigma, M):
    gx, gy, hx, hy = kernels(kw, sigma, M)
    Kx, Ky = orc.srsrk_kernel(kw, 1.0, sigma)
    c = kw // 2
    for spec, K in ((gx, Kx), (gy, Ky)):
        pad = np.zeros((M, M))
        idx = (np.arange(kw) - c) % M
        pad[np.ix_(idx, idx)] = K
        ref = np.fft.rfft2(pad)            # [k_x][k_y], k_y in [0, M/2]
        got = spec.cpu().numpy().astype(np.complex128)
        assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()


def test_precompute_kernel_freq_matches_numpy():
    torch = _torch()
    rng = np.random.default_rng(1)
    G = rng.standard_normal((9, 9)).astype(np.float32)
    out = srk.precompute_kernel_freq(torch.from_numpy(G), 256).cpu().numpy()
    pad = np.zeros((256, 256))
    idx = (np.arange(9) - 4) % 256
    pad[np.ix_(idx, idx)] = G
    ref = np.fft.rfft2(pad)
    assert np.abs(out - ref).max() <= 1e-5 * np.abs(ref).max()


# ---- forward -----------------------------------------------------------------

@pytest.mark.parametrize("M", [256, 512])
def test_forward_oas_small(orc, M):
    """adjoint/main_adjoint.cu Task 1 geometry with the SRSRK kernel: sigma=5,
    KW=31, erosion=16, W=256, N=288."""
    N, e, kw, sigma = 288, 16, 31, 5.0
    W = N - 2 * e
    gx, gy, Kx, Ky = kernels(kw, sigma, M)
    f = patterns.grf(N, 1000)[0]
    g = patterns.grf(N, 2000)[0]
    ws = srk.Workspace(N, W, e, M)
    R = dev(np.zeros((W, W)))
    ux, uy = dev(np.zeros((W, W))), dev(np.zeros((W, W)))
    srk.srkForwardOaS(dev(g), dev(f), gx, gy, N, W, e, 1.0, ws, R, ux, uy)
    ref = orc.forward(f, g, e, Kx, Ky)
    assert_close(host(R), ref, what="h")
    rux, ruy = orc.conv_valid(g, e, Kx, Ky)
    assert_close(host(ux), rux, what="ux")
    assert_close(host(uy), ruy, what="uy")


def test_training_forward_config0(orc):
    """Config 0: B=1..3, N=256, KW=31 (erosion 15, W=226), dx=1, E=I GRF; also E != I."""
    N, e, kw, sigma = 256, 15, 31, 5.0
    W = N - 2 * e
    gx, gy, Kx, Ky = kernels(kw, sigma, N)
    E = patterns.grf(N, 1000, B=3)
    I = patterns.grf(N, 2000, B=3)
    ws = srk.Workspace(N, W, e, N, B_max=3)
    R = dev(np.zeros((3, W, W)))
    dE = dev(E)
    srk.srkTrainingForward(dE, dE, gx, gy, N, W, e, 1.0, 3, ws, R)
    Rh = host(R)
    for b in range(3):
        assert_close(Rh[b], orc.forward(E[b], E[b], e, Kx, Ky), what=f"E=I b{b}")
    srk.srkTrainingForward(dE, dev(I), gx, gy, N, W, e, 1.0, 3, ws, R)
    Rh = host(R)
    for b in range(3):
        assert_close(Rh[b], orc.forward(I[b], E[b], e, Kx, Ky), what=f"E!=I b{b}")


def test_batch_demo_known_answer_gpu():
    """training/main_batch_demo.cu (sigma=20, KW=121, erosion 61, N=256, B=4 LNS)
    against the recorded stats of notebooks/plot_batch_demo.ipynb:225-228."""
    N, e = 256, 61
    W = N - 2 * e
    gx, gy, _, _ = kernels(121, 20.0, N)
    sig = [8.0, 6.0, 8.0, 6.0]
    E = np.stack([patterns.lns_pulse_train(N, s, vertical=b < 2) for b, s in enumerate(sig)])
    ws = srk.Workspace(N, W, e, N, B_max=4)
    R = dev(np.zeros((4, W, W)))
    dE = dev(E)
    srk.srkTrainingForward(dE, dE, gx, gy, N, W, e, 1.0, 4, ws, R)
    Rh = host(R)
    expect = [(-0.234, -6.484e-2, 8.204e-2), (-0.288, -1.029e-1, 1.074e-1)] * 2
    for b in range(4):
        mn, mean, std = expect[b]
        assert abs(Rh[b].min() - mn) < 5e-4
        assert Rh[b].max() <= 1e-5
        assert abs(Rh[b].mean() - mean) <= 0.5e-3 * 10 ** np.floor(np.log10(abs(mean))) * 1.001
        assert abs(Rh[b].std() - std) <= 0.5e-3 * 10 ** np.floor(np.log10(abs(std))) * 1.001


@pytest.mark.parametrize("M", [256, 512])
def test_forward_multi_tile_n1024(orc, M):
    """N = 1024, KW = 61 (erosion 30): 25 (M=256) / 9 (M=512) overlap-and-save tiles."""
    N, e, kw, sigma = 1024, 30, 61, 10.0
    W = N - 2 * e
    gx, gy, Kx, Ky = kernels(kw, sigma, M)
    f = patterns.grf(N, 1001)[0]
    g = patterns.grf(N, 2001)[0]
    ws = srk.Workspace(N, W, e, M, B_max=2)
    R = srk.forward_batched(dev(np.stack([g, f])), dev(np.stack([f, g])), gx, gy, e, 1.0, ws)
    Rh = host(R)
    assert_close(Rh[0], orc.forward(f, g, e, Kx, Ky, fft=True), what="b0")
    assert_close(Rh[1], orc.forward(g, f, e, Kx, Ky, fft=True), what="b1")


# ---- adjoints ----------------------------------------------------------------

@pytest.mark.parametrize("M", [256, 512])
def test_adjoint_recompute(orc, M):
    N, e, kw, sigma = 288, 16, 31, 5.0
    W = N - 2 * e
    gx, gy, Kx, Ky = kernels(kw, sigma, M)
    f = patterns.grf(N, 1002)[0]
    g = patterns.grf(N, 2002)[0]
    v = np.random.default_rng(123).standard_normal((W, W)).astype(np.float32)
    ws = srk.Workspace(N, W, e, M)
    lf, lg = srk.adjoint_batched(dev(v[None]), dev(g[None]), dev(f[None]), gx, gy, e, 1.0, ws)
    rf, rg = orc.adjoint(v, f, g, e, Kx, Ky)
    assert_close(host(lf)[0], rf, what="lambda_f")
    assert_close(host(lg)[0], rg, what="lambda_g")


def test_adjoint_srk_with_stored_u(orc):
    N, e, kw, sigma, M = 288, 16, 31, 5.0, 256
    W = N - 2 * e
    gx, gy, Kx, Ky = kernels(kw, sigma, M)
    f = patterns.grf(N, 1003)[0]
    g = patterns.grf(N, 2003)[0]
    v = np.random.default_rng(7).standard_normal((W, W)).astype(np.float32)
    ux, uy = orc.conv_valid(g, e, Kx, Ky)
    ws = srk.Workspace(N, W, e, M)
    lf, lg = dev(np.zeros((N, N))), dev(np.zeros((N, N)))
    srk.adjointSRK(dev(v), dev(ux), dev(uy), dev(f), gx, gy, N, W, e, 1.0, ws, lf, lg)
    rf, rg = orc.adjoint(v, f, g, e, Kx, Ky)
    assert_close(host(lf), rf, what="lambda_f")
    assert_close(host(lg), rg, what="lambda_g")


def test_adjoint_dot_product_gpu():
    """<h(df, g), v> = <df, lambda_f>, <h(f, dg), v> = <dg, lambda_g> on GPU results
    (SRSRK sigma=10, KW=61, N=512 whole-image tile)."""
    N, e, kw, sigma, M = 512, 30, 61, 10.0, 512
    W = N - 2 * e
    gx, gy, _, _ = kernels(kw, sigma, M)
    f, g, df, dg = (patterns.grf(N, s)[0] for s in (11, 12, 13, 14))
    v = np.random.default_rng(5).standard_normal((W, W)).astype(np.float32)
    ws = srk.Workspace(N, W, e, M)
    lf, lg = srk.adjoint_batched(dev(v[None]), dev(g[None]), dev(f[None]), gx, gy, e, 1.0, ws)
    h_df = host(srk.forward_batched(dev(g[None]), dev(df[None]), gx, gy, e, 1.0, ws))[0]
    h_dg = host(srk.forward_batched(dev(dg[None]), dev(f[None]), gx, gy, e, 1.0, ws))[0]
    a1, b1 = np.sum(h_df * v), np.sum(df * host(lf)[0])
    a2, b2 = np.sum(h_dg * v), np.sum(dg * host(lg)[0])
    assert abs(a1 - b1) <= 1e-4 * abs(a1)
    assert abs(a2 - b2) <= 1e-4 * abs(a2)


# ---- tangent / tangent-adjoint ------------------------------------------------

@pytest.mark.parametrize("M", [256, 512])
def test_tangent(orc, M):
    N, e, kw, sigma = 288, 16, 31, 5.0
    W = N - 2 * e
    gx, gy, Kx, Ky = kernels(kw, sigma, M)
    f, g, df, dg = (patterns.grf(N, s)[0] for s in (42, 43, 100, 101))
    ws = srk.Workspace(N, W, e, M)
    dh = dev(np.zeros((W, W)))
    srk.shrinkageTangentOaS(dev(f), dev(g), dev(df), dev(dg), gx, gy, N, W, e, 1.0, ws, dh)
    assert_close(host(dh), orc.tangent(f, g, df, dg, e, Kx, Ky), what="dh")


@pytest.mark.parametrize("M", [256, 512])
def test_tangent_adjoint(orc, M):
    N, e, kw, sigma = 288, 16, 31, 5.0
    W = N - 2 * e
    gx, gy, Kx, Ky = kernels(kw, sigma, M)
    f, g, df, dg = (patterns.grf(N, s)[0] for s in (42, 43, 103, 104))
    rng = np.random.default_rng(200)
    lam = rng.standard_normal((W, W)).astype(np.float32)
    dlam = rng.standard_normal((W, W)).astype(np.float32)
    ws = srk.Workspace(N, W, e, M)
    a, b = dev(np.zeros((N, N))), dev(np.zeros((N, N)))
    srk.shrinkageTangentAdjointOaS(dev(f), dev(g), dev(df), dev(dg), dev(lam), dev(dlam), gx, gy,
                                   N, W, e, 1.0, ws, a, b)
    ra, rb = orc.tangent_adjoint(f, g, df, dg, lam, dlam, e, Kx, Ky)
    assert_close(host(a), ra, what="dlam_f")
    assert_close(host(b), rb, what="dlam_g")


def test_adjoint_multi_tile_n1024(orc):
    N, e, kw, sigma, M = 1024, 30, 61, 10.0, 256
    W = N - 2 * e
    gx, gy, Kx, Ky = kernels(kw, sigma, M)
    f = patterns.grf(N, 1004)[0]
    g = patterns.grf(N, 2004)[0]
    v = np.random.default_rng(9).standard_normal((W, W)).astype(np.float32)
    ws = srk.Workspace(N, W, e, M)
    lf, lg = srk.adjoint_batched(dev(v[None]), dev(g[None]), dev(f[None]), gx, gy, e, 1.0, ws)
    rf, rg = orc.adjoint(v, f, g, e, Kx, Ky, fft=True)
    assert_close(host(lf)[0], rf, what="lambda_f")
    assert_close(host(lg)[0], rg, what="lambda_g")
