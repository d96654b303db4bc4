"""The headline batch at full size: srkTrainingForward + srkAdjointBatched over
N = 512, B = 5000, KW = 121 (erosion 60, W = 392), E = I, as the training
contract calls them (training/main_cd_training.cu:93-121,
training/srk_training_forward.cuh:175-241; adjoint/srk_adjoint.cuh:402-425).

At this size the batch runs as many chunks of tiles alternating between two
streams, so the checks are ones that hold for every image whatever chunk and
stream it lands in:
- sampled images (first, last, both sides of the forward and adjoint chunk
  boundaries) equal, bit for bit, what a one-image workspace computes;
- two sampled images match the FP64 oracle at the fp32 tolerance;
- every image satisfies the bilinear identities <h, v> = <f, lambda_f> =
  <g, lambda_g> (h is linear in f for fixed g and in g for fixed f).
"""
import numpy as np
import pytest

import shrinkage_full_tests_amd as srk
from shrinkage_full_tests_amd import patterns

pytestmark = pytest.mark.gpu

N, KW, SIGMA, B = 512, 121, 20.0, 5000
E_ = (KW - 1) // 2
W = N - 2 * E_
TOL_REL, TOL_ABS = 1e-3, 1e-5
# |<h,v> - <f,lambda>| against the norm products: the outputs carry FP32 rounding
# (about 1e-7 relative per element); an image taken from the wrong chunk slot or
# a sign error moves the difference to order 1e-2 and beyond
ID_TOL = 1e-5


def _run(torch, ws, E, v):
    R = torch.empty((E.shape[0], W, W), device="cuda")
    lf = torch.empty((E.shape[0], N, N), device="cuda")
    lg = torch.empty((E.shape[0], N, N), device="cuda")
    gx, gy = srk.kernel_spectra_srsrk(KW, SIGMA, N)
    srk.srkTrainingForward(E, E, gx, gy, N, W, E_, 1.0, E.shape[0], ws, R)
    srk.adjoint_batched(v, E, E, gx, gy, E_, 1.0, ws, lf, lg)
    torch.cuda.synchronize()
    return R, lf, lg


@pytest.fixture(scope="module")
def batch():
    import torch
    E = patterns.grf_torch(B, N, seed=51)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(52)
    v = torch.randn((B, W, W), generator=gen, device="cuda")
    return E, v


@pytest.mark.parametrize("prec", ["fp32", "bf16_exchange"])
def test_full_batch_images_match_single_image_runs(batch, prec):
    import torch
    E, v = batch
    ws = srk.Workspace(N, W, E_, N, B_max=B, precision=prec)
    info = ws.info()
    R, lf, lg = _run(torch, ws, E, v)
    cf, ca = info["chunk_fwd"], info["chunk_adj"]
    assert cf < B and ca < B, "the batch must span several chunks"
    picks = sorted({0, 1, cf - 1, cf, ca - 1, ca, 2 * ca, B // 2, B - 2, B - 1})
    one = srk.Workspace(N, W, E_, N, B_max=1, precision=prec)
    for b in picks:
        r1, f1, g1 = _run(torch, one, E[b:b + 1].contiguous(), v[b:b + 1].contiguous())
        for name, got, want in (("h", R[b], r1[0]), ("lambda_f", lf[b], f1[0]),
                                ("lambda_g", lg[b], g1[0])):
            assert torch.equal(got, want), f"{prec}: image {b} {name} differs from its B = 1 run"


def test_full_batch_oracle_and_bilinear_identities(orc, batch):
    import torch
    E, v = batch
    ws = srk.Workspace(N, W, E_, N, B_max=B)
    R, lf, lg = _run(torch, ws, E, v)
    gx, gy, hx, hy = srk.kernel_spectra_srsrk(KW, SIGMA, N, with_spatial=True)
    Kx, Ky = hx.numpy().astype(np.float64), hy.numpy().astype(np.float64)
    for b in (7, B - 3):
        f = E[b].cpu().numpy().astype(np.float64)
        vv = v[b].cpu().numpy().astype(np.float64)
        ref_h = orc.forward(f, f, E_, Kx, Ky, fft=True)
        ref_f, ref_g = orc.adjoint(vv, f, f, E_, Kx, Ky, fft=True)
        for name, got, ref in (("h", R[b], ref_h), ("lambda_f", lf[b], ref_f),
                               ("lambda_g", lg[b], ref_g)):
            a = got.cpu().numpy().astype(np.float64)
            err = np.abs(a - ref)
            lim = np.maximum(TOL_ABS, TOL_REL * np.maximum(np.abs(a), np.abs(ref)))
            assert not (err > lim).any(), f"image {b} {name}: max err {err.max():.3e}"
    # per-image inner products in float64 on the device
    Rd, vd, Ed = R.double(), v.double(), E.double()
    hv = (Rd * vd).sum(dim=(1, 2))
    nrv = Rd.flatten(1).norm(dim=1) * vd.flatten(1).norm(dim=1)
    for name, lam in (("lambda_f", lf), ("lambda_g", lg)):
        ld = lam.double()
        fl = (Ed * ld).sum(dim=(1, 2))
        scale = nrv + Ed.flatten(1).norm(dim=1) * ld.flatten(1).norm(dim=1)
        ratio = ((hv - fl).abs() / scale).max().item()
        assert ratio < ID_TOL, f"<h,v> = <f,{name}> off by {ratio:.3e} of the norm products"
