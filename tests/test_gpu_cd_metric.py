"""GPU CD metric (srkMeasureCDBatched, srkShrinkageCDBatched) against the FP64
host restatement of MeasureCD_Batched (third_party/src/metricUtil/
cd_measure_util.cu:310-446) and of the crop / E + c R / centre-line steps of
training/main_cd_training.cu:128-165.  Pass: |CD_gpu - CD_ref| < 0.01 px
(README.md:36); the edges themselves are compared to the same bound."""
import numpy as np
import pytest

import shrinkage_full_tests_amd as srk
from oracle import cd as cdm
from shrinkage_full_tests_amd import patterns

pytestmark = pytest.mark.gpu

TOL_PX = 0.01


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy().astype(np.float64)


def _check(got, sig, L, A, thr):
    # the reference sample is FP32: compare with the oracle on the FP32 values
    ref = cdm.measure_cd(np.asarray(sig, np.float32).astype(np.float64), L, A, thr)
    for g, r, name in zip(got, ref, ("left", "right", "cd")):
        g = host(g)
        fin = np.isfinite(r)
        assert np.array_equal(np.isfinite(g), fin), (name, g, r)
        assert np.all(g[~fin] == r[~fin]), (name, g, r)
        assert np.abs(g[fin] - r[fin]).max(initial=0.0) < TOL_PX, (name, g, r)
    return ref


@pytest.mark.parametrize("N,L,A", [(136, 16, 8), (392, 16, 8), (1000, 16, 8), (257, 4, 3)])
def test_measure_cd_smooth_signals(N, L, A):
    rng = np.random.default_rng(N + L)
    x = np.arange(N)
    sigs = []
    for b in range(6):
        w = rng.uniform(0.15, 0.45) * N
        c = N / 2 + rng.uniform(-0.1, 0.1) * N
        s = rng.uniform(1.5, 4.0)
        line = 1.0 / (1.0 + np.exp(-(x - (c - w / 2)) / s)) - 1.0 / (1.0 + np.exp(-(x - (c + w / 2)) / s))
        if b % 2:
            line = 1.0 - line           # trench instead of line
        sigs.append(line + 0.02 * rng.standard_normal(N))
    sig = np.stack(sigs)
    ref = _check(srk.measure_cd_batched(dev(sig), L, A, threshold=0.5), sig, L, A, 0.5)
    assert np.all(np.isfinite(ref[2]))


def test_measure_cd_gaussian_fwhm():
    """Closed form: a Gaussian of std s crosses half its peak at +-s sqrt(2 ln 2)."""
    N, s = 300, 9.0
    x = np.arange(N) - (N - 1) / 2
    sig = np.exp(-0.5 * (x / s) ** 2)[None]
    _, _, cd = srk.measure_cd_batched(dev(sig), threshold=0.5)
    assert abs(host(cd)[0] - 2 * s * np.sqrt(2 * np.log(2))) < 0.02


def test_measure_cd_missing_edges_and_zero_threshold():
    N = 128
    x = np.arange(N)
    sig = np.stack([np.ones(N),                          # no crossing at all
                    np.where(x < 80, -1.0, 1.0),         # right edge only
                    np.sin(2 * np.pi * (x - 64) / 50.0)])  # zero crossings (threshold 0)
    _check(srk.measure_cd_batched(dev(sig), 16, 8, threshold=0.0), sig, 16, 8, 0.0)


def test_measure_cd_rejects_oversized_signals():
    import torch
    sig = torch.zeros((1, 50000), device="cuda")
    with pytest.raises(srk.SrkError):
        srk.measure_cd_batched(sig)


@pytest.mark.parametrize("kw", [61, 121])
def test_shrinkage_cd_fused_matches_host_pipeline(orc, kw):
    """Forward on LNS pulse trains (reference Sections B/C), then the fused CD
    on the device against (a) the host pipeline on the same GPU R, (b) the
    FP64 oracle's R."""
    import torch
    N, e, c = 256, (kw - 1) // 2, 1.0
    W = N - 2 * e
    sigs = [6.0, 8.0, 11.0]
    imgs = np.stack([patterns.lns_pulse_train(N, s, vertical=True) for s in sigs] +
                    [patterns.lns_pulse_train(N, s, vertical=False) for s in sigs])
    B = imgs.shape[0]
    gx, gy, hx, hy = srk.kernel_spectra_srsrk(kw, e / 3.0, N, with_spatial=True)
    Kx, Ky = hx.numpy().astype(np.float64), hy.numpy().astype(np.float64)
    ws = srk.Workspace(N, W, e, N, B_max=B)
    dE = dev(imgs)
    R = torch.empty((B, W, W), device="cuda")
    srk.srkTrainingForward(dE, dE, gx, gy, N, W, e, 1.0, B, ws, R)
    half = B // 2
    got_v = srk.shrinkage_cd_batched(dE[:half], R[:half], e, c, center_row=True)
    got_h = srk.shrinkage_cd_batched(dE[half:], R[half:], e, c, center_row=False)
    cd_gpu = np.concatenate([host(got_v[2]), host(got_h[2])])
    Rg = host(R)
    sig_same, sig_ref = [], []
    for b in range(B):
        E = imgs[b].astype(np.float64)
        crop = E[e:N - e, e:N - e]
        sg = (crop.astype(np.float32) + np.float32(c) * Rg[b].astype(np.float32)).astype(np.float64)
        sr = crop + c * orc.forward(E, E, e, Kx, Ky, fft=True)
        # arrays are [column, row]: the centre row is [:, W // 2]
        sig_same.append(sg[:, W // 2] if b < half else sg[W // 2, :])
        sig_ref.append(sr[:, W // 2] if b < half else sr[W // 2, :])
    _, _, cd_same = cdm.measure_cd(np.stack(sig_same))
    _, _, cd_ref = cdm.measure_cd(np.stack(sig_ref))
    assert np.all(np.isfinite(cd_ref))
    assert np.abs(cd_gpu - cd_same).max() < TOL_PX, (cd_gpu, cd_same)
    assert np.abs(cd_gpu - cd_ref).max() < TOL_PX, (cd_gpu, cd_ref)
