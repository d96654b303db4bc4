"""The fused one-launch pipeline against the three-pass path and the FP64 oracle.

FP32 forward and adjoint calls on one-tile images (N <= M, M = 256 or 512)
run as one persistent launch when SRK_TUNE_FUSED is 1 (the default is the
three-pass path, which is faster at every measured size).  Each case runs
both forms on one workspace, checks that no wait of the fused launch passed
its deadline, compares the two forms on every image (same transforms, same
order of operations per line: expected equal up to FP32 rounding of the
band-height-dependent scheduling, checked at 1e-5 relative), and compares the
fused form with the oracle on the first and last image at the FP32 tolerance
(adjoint/main_adjoint.cu:70-90 pass criterion).
"""
import numpy as np
import pytest

import shrinkage_full_tests_amd as srk
from shrinkage_full_tests_amd import patterns

from test_gpu_parity import assert_close, host

pytestmark = pytest.mark.gpu

# N, KW, M, B: full XCD rounds (B = 16: two images per XCD), one image, odd
# batches that leave XCDs idle, N < M with a ragged column count, and a batch
# that takes many pipeline rounds per XCD
CASES = [(512, 121, 512, 16), (512, 121, 512, 1), (256, 31, 256, 13), (200, 31, 256, 9),
         (300, 61, 512, 11), (256, 61, 256, 203)]


def _run(ws, E, I, v, gx, gy, e, fused):
    import torch
    ws.set_tuning(fused=fused)
    h = srk.forward_batched(E, I, gx, gy, e, 1.0, ws)
    lf, lg = srk.adjoint_batched(v, E, I, gx, gy, e, 1.0, ws)
    torch.cuda.synchronize()
    return h, lf, lg


@pytest.mark.parametrize("N,kw,M,B", CASES)
def test_fused_matches_three_pass_and_oracle(orc, N, kw, M, B):
    import torch
    print(f"\n[fused] N={N} KW={kw} M={M} B={B}", flush=True)
    e = (kw - 1) // 2
    W = N - 2 * e
    gx, gy, hx, hy = srk.kernel_spectra_srsrk(kw, e / 3.0, M, with_spatial=True)
    E = patterns.grf_torch(B, N, seed=11)
    I = patterns.grf_torch(B, N, seed=12)
    v = torch.randn((B, W, W), device="cuda", generator=torch.Generator("cuda").manual_seed(13))
    ws = srk.Workspace(N, W, e, M, B_max=B)
    available, _ = ws.fused_status()
    assert available
    ref3 = _run(ws, E, I, v, gx, gy, e, 0)
    got = _run(ws, E, I, v, gx, gy, e, 1)
    _, timeouts = ws.fused_status()
    assert timeouts == 0
    for name, a, b in zip(("h", "lambda_f", "lambda_g"), got, ref3):
        scale = b.abs().max().item()
        d = (a - b).abs().max().item()
        print(f"  {name}: max|fused - three-pass| {d:.3e} (max {scale:.3e})", flush=True)
        assert d <= 1e-5 * scale, name
    Kx, Ky = hx.numpy().astype(np.float64), hy.numpy().astype(np.float64)
    for b in sorted({0, B - 1}):
        f_, g_ = host(I[b]), host(E[b])
        ref_h = orc.forward(f_, g_, e, Kx, Ky, fft=True)
        ref_f, ref_g = orc.adjoint(host(v[b]), f_, g_, e, Kx, Ky, fft=True)
        assert_close(host(got[0][b]), ref_h, what=f"h b={b}")
        assert_close(host(got[1][b]), ref_f, what=f"lambda_f b={b}")
        assert_close(host(got[2][b]), ref_g, what=f"lambda_g b={b}")


def test_fused_is_used_and_repeatable():
    """Two fused calls in a row give bit-equal results; the kernel timer sees
    one fused launch per call and no three-pass launch."""
    import torch
    N, kw, M, B = 512, 121, 512, 24
    e = (kw - 1) // 2
    W = N - 2 * e
    print(f"\n[fused repeat] N={N} M={M} B={B}", flush=True)
    gx, gy = srk.kernel_spectra_srsrk(kw, e / 3.0, M)
    E = patterns.grf_torch(B, N, seed=21)
    I = patterns.grf_torch(B, N, seed=22)
    ws = srk.Workspace(N, W, e, M, B_max=B)
    ws.set_tuning(fused=1)
    ws.set_timing(True)
    h1 = srk.forward_batched(E, I, gx, gy, e, 1.0, ws)
    h2 = srk.forward_batched(E, I, gx, gy, e, 1.0, ws)
    torch.cuda.synchronize()
    kt = ws.kernel_times()
    assert torch.equal(h1, h2)
    assert kt["k_fused"][1] == 2
    assert all(kt[k][1] == 0 for k in kt if k != "k_fused")
    assert ws.fused_status() == (True, 0)


@pytest.mark.parametrize("slots,lag", [(2, 1), (3, 2), (4, 3), (4, 1)])
def test_fused_ring_shapes(slots, lag):
    """Every exchange-ring depth and stage lag gives the three-pass result."""
    import torch
    N, kw, M, B = 256, 31, 256, 40
    e = (kw - 1) // 2
    W = N - 2 * e
    print(f"\n[fused ring] slots={slots} lag={lag}", flush=True)
    gx, gy = srk.kernel_spectra_srsrk(kw, e / 3.0, M)
    E = patterns.grf_torch(B, N, seed=31)
    I = patterns.grf_torch(B, N, seed=32)
    v = torch.randn((B, W, W), device="cuda", generator=torch.Generator("cuda").manual_seed(33))
    ws = srk.Workspace(N, W, e, M, B_max=B)
    ref = _run(ws, E, I, v, gx, gy, e, 0)
    ws.set_tuning(fused_slots=slots, fused_lag=lag)
    got = _run(ws, E, I, v, gx, gy, e, 1)
    assert ws.fused_status() == (True, 0)
    for a, b in zip(got, ref):
        assert (a - b).abs().max().item() <= 1e-5 * b.abs().max().item()


def test_fused_ring_limits():
    N, kw, M = 256, 31, 256
    e = (kw - 1) // 2
    ws = srk.Workspace(N, N - 2 * e, e, M)
    with pytest.raises(srk.SrkError):
        ws.set_tuning(fused_slots=5)
    with pytest.raises(srk.SrkError):
        ws.set_tuning(fused_lag=4)   # not below the 4 slots
    with pytest.raises(srk.SrkError):
        ws.set_tuning(fused_lag=0)
    ws.set_tuning(fused_slots=2, fused_lag=1)
    with pytest.raises(srk.SrkError):
        ws.set_tuning(fused_slots=1)  # not above the lag


def test_fused_single_image_with_u_outputs():
    """srkForwardOaS through the fused launch (one-tile image) writes h, u_x and
    u_y equal to the three-pass path's."""
    import torch
    N, kw, M = 300, 61, 512
    e = (kw - 1) // 2
    W = N - 2 * e
    print(f"\n[fused u] N={N} M={M}", flush=True)
    gx, gy = srk.kernel_spectra_srsrk(kw, e / 3.0, M)
    E = patterns.grf_torch(1, N, seed=41)[0].contiguous()
    I = patterns.grf_torch(1, N, seed=42)[0].contiguous()
    ws = srk.Workspace(N, W, e, M)
    out = {}
    for fused in (0, 1):
        ws.set_tuning(fused=fused)
        R, ux, uy = (torch.full((W, W), float("nan"), device="cuda") for _ in range(3))
        srk.srkForwardOaS(E, I, gx, gy, N, W, e, 1.0, ws, R, ux, uy)
        torch.cuda.synchronize()
        out[fused] = (R, ux, uy)
    assert ws.fused_status() == (True, 0)
    for a, b in zip(out[1], out[0]):
        assert not torch.isnan(a).any()
        assert (a - b).abs().max().item() <= 1e-5 * b.abs().max().item()


def test_fused_graph_replays_match_eager():
    """The fused launch (its argument-setup kernel and the persistent kernel)
    captured in a hipGraph: several replays, with eager work between them,
    each equal to the eager result."""
    import torch
    N, kw, M, B = 256, 31, 256, 24
    e = (kw - 1) // 2
    W = N - 2 * e
    print(f"\n[fused graph] N={N} M={M} B={B}", flush=True)
    gx, gy = srk.kernel_spectra_srsrk(kw, e / 3.0, M)
    E = patterns.grf_torch(B, N, seed=51)
    v = torch.randn((B, W, W), device="cuda", generator=torch.Generator("cuda").manual_seed(52))
    ws = srk.Workspace(N, W, e, M, B_max=B)
    ws.set_tuning(fused=1)
    R = torch.empty((B, W, W), device="cuda")
    lf = torch.empty((B, N, N), device="cuda")
    lg = torch.empty((B, N, N), device="cuda")

    def step():
        srk.forward_batched(E, E, gx, gy, e, 1.0, ws, R)
        srk.adjoint_batched(v, E, E, gx, gy, e, 1.0, ws, lf, lg)

    step()
    torch.cuda.synchronize()
    ref = [R.clone(), lf.clone(), lg.clone()]
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            step()
    torch.cuda.current_stream().wait_stream(s)
    for it in range(3):
        R.zero_(); lf.zero_(); lg.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for got, want in zip((R, lf, lg), ref):
            assert torch.equal(got, want), f"replay {it} differs from eager"
        _ = (E * 2).sum().item()   # eager work between replays
    assert ws.fused_status() == (True, 0)
