"""Reduced-precision forms against the FP64 oracle at the fp16 tolerance of
README.md:8-10 (tol_rel 1e-2, tol_abs 1e-3), plus CD parity on line/space
patterns (< 0.01 px): "mixed_fp16" (FP16 MFMA DFT stages, FP32 accumulation)
and "bf16_exchange" (FP32 transforms, bfloat16 spectra between the passes)."""
import numpy as np
import pytest

import shrinkage_full_tests_amd as srk
from oracle import cd as cdm
from shrinkage_full_tests_amd import patterns

pytestmark = pytest.mark.gpu

TOL_REL, TOL_ABS = 1e-2, 1e-3
PRECS = ["mixed_fp16", "bf16_exchange"]


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def assert_close(a, n, what=""):
    err = np.abs(a - n)
    lim = np.maximum(TOL_ABS, TOL_REL * np.maximum(np.abs(a), np.abs(n)))
    bad = err > lim
    assert not bad.any(), f"{what}: {bad.sum()} violations, max err {err.max():.3e}, max|ref| {np.abs(n).max():.3e}"


def kernels(kw, sigma, M):
    gx, gy, hx, hy = srk.kernel_spectra_srsrk(kw, sigma, M, with_spatial=True)
    return gx, gy, hx.numpy().astype(np.float64), hy.numpy().astype(np.float64)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("N,kw,M", [(256, 31, 256), (288, 31, 256), (512, 121, 512), (288, 31, 512),
                                    (1024, 61, 1024), (1500, 61, 512)])
def test_mixed_forward_adjoint(orc, N, kw, M, prec):
    e = (kw - 1) // 2 + (1 if N == 288 else 0)
    W = N - 2 * e
    sigma = ((kw - 1) // 2) / 3.0
    gx, gy, Kx, Ky = kernels(kw, sigma, M)
    f = patterns.grf(N, 1000 + N)[0]
    g = patterns.grf(N, 2000 + N)[0]
    v = np.random.default_rng(N).standard_normal((W, W)).astype(np.float32)
    ws = srk.Workspace(N, W, e, M, precision=prec)
    h = srk.forward_batched(dev(g[None]), dev(f[None]), gx, gy, e, 1.0, ws)
    lf, lg = srk.adjoint_batched(dev(v[None]), dev(g[None]), dev(f[None]), gx, gy, e, 1.0, ws)
    fft = N > 600
    assert_close(host(h)[0], orc.forward(f, g, e, Kx, Ky, fft=fft), "h")
    rf, rg = orc.adjoint(v, f, g, e, Kx, Ky, fft=fft)
    assert_close(host(lf)[0], rf, "lambda_f")
    assert_close(host(lg)[0], rg, "lambda_g")


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("M", [256, 512])
def test_mixed_tangent_and_tangent_adjoint(orc, M, prec):
    N, e, kw, sigma = 288, 16, 31, 5.0
    W = N - 2 * e
    gx, gy, Kx, Ky = kernels(kw, sigma, M)
    f, g, df, dg = (patterns.grf(N, s)[0] for s in (42, 43, 100, 101))
    rng = np.random.default_rng(200)
    lam = rng.standard_normal((W, W)).astype(np.float32)
    dlam = rng.standard_normal((W, W)).astype(np.float32)
    ws = srk.Workspace(N, W, e, M, precision=prec)
    dh = dev(np.zeros((W, W)))
    srk.shrinkageTangentOaS(dev(f), dev(g), dev(df), dev(dg), gx, gy, N, W, e, 1.0, ws, dh)
    assert_close(host(dh), orc.tangent(f, g, df, dg, e, Kx, Ky), "dh")
    a, b = dev(np.zeros((N, N))), dev(np.zeros((N, N)))
    srk.shrinkageTangentAdjointOaS(dev(f), dev(g), dev(df), dev(dg), dev(lam), dev(dlam), gx, gy,
                                   N, W, e, 1.0, ws, a, b)
    ra, rb = orc.tangent_adjoint(f, g, df, dg, lam, dlam, e, Kx, Ky)
    assert_close(host(a), ra, "dlam_f")
    assert_close(host(b), rb, "dlam_g")


@pytest.mark.parametrize("prec", PRECS)
def test_mixed_cd_box_lines(orc, prec):
    import torch
    N, kw = 256, 121
    e = 60
    W = N - 2 * e
    imgs = np.stack([patterns.lns_box(N, vertical=True), patterns.lns_box(N, vertical=False),
                     patterns.lns_box(N, duty=0.4, vertical=True)])
    vert = [True, False, True]
    gx, gy, Kx, Ky = kernels(kw, 20.0, N)
    ws = srk.Workspace(N, W, e, N, B_max=3, precision=prec)
    R = torch.empty((3, W, W), device="cuda")
    dE = dev(imgs)
    srk.srkTrainingForward(dE, dE, gx, gy, N, W, e, 1.0, 3, ws, R)
    Rg = host(R)
    sg, sr = [], []
    for b in range(3):
        E = imgs[b].astype(np.float64)
        Rr = orc.forward(E, E, e, Kx, Ky, fft=True)
        crop = E[e:N - e, e:N - e]
        a, r = crop + Rg[b], crop + Rr
        sg.append(a[:, W // 2] if vert[b] else a[W // 2, :])
        sr.append(r[:, W // 2] if vert[b] else r[W // 2, :])
    _, _, cg = cdm.measure_cd(np.stack(sg), threshold=0.5)
    _, _, cr = cdm.measure_cd(np.stack(sr), threshold=0.5)
    assert np.all(cdm.found(cr)) and np.abs(cg - cr).max() < 0.01, (cg, cr)


def test_mixed_precision_rejected_for_large_tiles():
    with pytest.raises(srk.SrkError):
        srk.Workspace(4096, 4096 - 60, 30, 2048, precision="mixed_fp16")


def test_bf16_exchange_large_tile_and_stored_u(orc):
    """M = 2048 tiles (not available to the MFMA form), and the single-image
    pair srkForwardOaS (u stored, computed in FP32) + adjointSRK (lambda_f from
    the stored u, lambda_g through the bfloat16 correlation plane)."""
    N, kw, M = 2300, 61, 2048
    e = 30
    W = N - 2 * e
    gx, gy, Kx, Ky = kernels(kw, 10.0, M)
    f = patterns.grf(N, 77)[0]
    g = patterns.grf(N, 78)[0]
    v = np.random.default_rng(79).standard_normal((W, W)).astype(np.float32)
    ws = srk.Workspace(N, W, e, M, precision="bf16_exchange")
    h = srk.forward_batched(dev(g[None]), dev(f[None]), gx, gy, e, 1.0, ws)
    assert_close(host(h)[0], orc.forward(f, g, e, Kx, Ky, fft=True), "h")
    rf, rg = orc.adjoint(v, f, g, e, Kx, Ky, fft=True)
    lf, lg = srk.adjoint_batched(dev(v[None]), dev(g[None]), dev(f[None]), gx, gy, e, 1.0, ws)
    assert_close(host(lf)[0], rf, "lambda_f")
    assert_close(host(lg)[0], rg, "lambda_g")
    ux = dev(np.zeros((W, W))); uy = dev(np.zeros((W, W))); R = dev(np.zeros((W, W)))
    srk.srkForwardOaS(dev(g), dev(f), gx, gy, N, W, e, 1.0, ws, R, ux, uy)
    assert_close(host(R), orc.forward(f, g, e, Kx, Ky, fft=True), "h (u stored)")
    lf2, lg2 = dev(np.zeros((N, N))), dev(np.zeros((N, N)))
    srk.adjointSRK(dev(v), ux, uy, dev(f), gx, gy, N, W, e, 1.0, ws, lf2, lg2)
    assert_close(host(lf2), rf, "lambda_f (stored u)")
    assert_close(host(lg2), rg, "lambda_g (stored u)")
