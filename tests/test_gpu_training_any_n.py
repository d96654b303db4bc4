"""srkTrainingForward as a drop-in at the sizes the reference's own tests use
(GPU), with spectra at size N and tiles of another size M.

- N = 192, B = 3000, sigma 10 (KW = 61, erosion 30, W = 132): the big
  batched configuration of training/main_forward_training.cu:413-420,507-599,
  with its Gaussian-process inputs (GpGenerator, length scale 24, seeds 60 /
  61, one image repeated over the batch), on M = 256 tiles.
- N = 288, erosion 16, KW = 31 (adjoint/main_adjoint.cu:249-254 geometry)
  on M = 512 (one tile) and M = 256 (overlap-and-save tiles), and an odd
  N = 193.
- Kernels whose support exceeds 2 erosion + 1 are rejected (SRK_ERR_KERNEL).
Results are compared with the FP64 oracle at the fp32 tolerance.
"""
import numpy as np
import pytest

import shrinkage_full_tests_amd as srk
from shrinkage_full_tests_amd import patterns

pytestmark = pytest.mark.gpu

TOL_REL, TOL_ABS = 1e-3, 1e-5


def assert_close(a, n, what=""):
    err = np.abs(a - n)
    lim = np.maximum(TOL_ABS, TOL_REL * np.maximum(np.abs(a), np.abs(n)))
    bad = err > lim
    assert not bad.any(), (f"{what}: {bad.sum()} violations, max err {err.max():.3e}, "
                           f"max|ref| {np.abs(n).max():.3e}")


def spectra_at(N, kw, sigma):
    """Spectra at size N from the FP32 spatial kernels, as the reference's
    precompute_G_freq_host does, and the FP64 kernels for the oracle."""
    import torch
    _, _, hx, hy = srk.kernel_spectra_srsrk(kw, sigma, 256 if kw <= 256 else 512,
                                            with_spatial=True)
    gx = srk.precompute_kernel_freq(hx, N)
    gy = srk.precompute_kernel_freq(hy, N)
    torch.cuda.synchronize()
    return gx, gy, hx.numpy().astype(np.float64), hy.numpy().astype(np.float64)


def test_reference_big_batch_n192(orc):
    import torch
    N, B, sigma, M = 192, 3000, 10.0, 256
    e = 30
    kw, W = 2 * e + 1, N - 2 * e
    gx, gy, Kx, Ky = spectra_at(N, kw, sigma)
    E0 = patterns.gp_field(N, 24.0, 60)
    I0 = patterns.gp_field(N, 24.0, 61)
    E = torch.from_numpy(E0).cuda()[None].repeat(B, 1, 1).contiguous()
    I = torch.from_numpy(I0).cuda()[None].repeat(B, 1, 1).contiguous()
    ws = srk.Workspace(N, W, e, M, B_max=B)
    R = torch.empty((B, W, W), device="cuda")
    srk.srkTrainingForward(E, I, gx, gy, N, W, e, 1.0, B, ws, R)
    torch.cuda.synchronize()
    ref = orc.forward(I0, E0, e, Kx, Ky)
    for b in (0, 1, B // 2, B - 1):
        assert_close(R[b].double().cpu().numpy(), ref, f"b={b}")
    assert bool((R == R[0:1]).all()), "identical images must give identical results"


@pytest.mark.parametrize("N,e,kw,M", [(288, 16, 31, 512), (288, 16, 31, 256), (193, 15, 31, 256),
                                      (300, 30, 61, 1024)])
def test_training_forward_other_sizes(orc, N, e, kw, M):
    import torch
    B = 3
    W = N - 2 * e
    gx, gy, Kx, Ky = spectra_at(N, kw, (kw - 1) / 6.0)
    E = patterns.grf(N, 910 + N, B=B)
    I = patterns.grf(N, 920 + N, B=B)
    ws = srk.Workspace(N, W, e, M, B_max=B)
    R = torch.empty((B, W, W), device="cuda")
    dE, dI = torch.from_numpy(E).cuda(), torch.from_numpy(I).cuda()
    for dx in (1.0, 4.0):
        srk.srkTrainingForward(dE, dI, gx, gy, N, W, e, dx, B, ws, R)
        Rh = R.double().cpu().numpy()
        for b in range(B):
            assert_close(Rh[b], orc.forward(I[b], E[b], e, Kx, Ky, s=dx * dx),
                         f"N={N} M={M} dx={dx} b={b}")
    # E = I, the CD-test contract
    srk.srkTrainingForward(dE, dE, gx, gy, N, W, e, 1.0, B, ws, R)
    Rh = R.double().cpu().numpy()
    assert_close(Rh[1], orc.forward(E[1], E[1], e, Kx, Ky), f"E=I N={N} M={M}")


def test_support_beyond_erosion_is_rejected():
    import torch
    N, M, kw = 192, 256, 61
    gx, gy, _, _ = spectra_at(N, kw, 10.0)
    # erosion 30 covers the kernel radius: accepted, tiny outside measure
    ws = srk.Workspace(N, N - 60, 30, M, B_max=1)
    r = srk.check_spectra_support(gx, gy, N, 30, ws)
    assert 0.0 <= r < 1e-6, r
    # erosion 20 < (KW - 1) / 2: the call fails instead of computing
    ws20 = srk.Workspace(N, N - 40, 20, M, B_max=1)
    E = torch.rand((1, N, N), device="cuda")
    R = torch.zeros((1, N - 40, N - 40), device="cuda")
    with pytest.raises(srk.SrkError) as ei:
        srk.srkTrainingForward(E, E, gx, gy, N, N - 40, 20, 1.0, 1, ws20, R)
    assert ei.value.status == -9
    with pytest.raises(srk.SrkError):
        srk.check_spectra_support(gx, gy, N, 20, ws20)
    # the same check on tile-size spectra (the M == N path of every entry point)
    gx256, gy256, _, _ = spectra_at(M, kw, 10.0)
    assert srk.check_spectra_support(gx256, gy256, M, 30, ws) < 1e-6
    with pytest.raises(srk.SrkError):
        srk.check_spectra_support(gx256, gy256, M, 20, ws20)
