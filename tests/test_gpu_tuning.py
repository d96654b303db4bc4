"""Every tuning variant of the engine gives the same operator: row-pass band
heights 16, 8, 4 and small launch chunks, in every precision form, on whole
image tiles (N = M) and on overlap-and-save tiles (N > M), against the FP64
oracle at the tolerance of each form (README.md:8-10)."""
import numpy as np
import pytest

import shrinkage_full_tests_amd as srk
from shrinkage_full_tests_amd import patterns

pytestmark = pytest.mark.gpu

TOL = {"fp32": (1e-3, 1e-5), "mixed_fp16": (1e-2, 1e-3), "bf16_exchange": (1e-2, 1e-3)}


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def assert_close(a, n, prec, what):
    tr, ta = TOL[prec]
    err = np.abs(a - n)
    lim = np.maximum(ta, tr * np.maximum(np.abs(a), np.abs(n)))
    bad = err > lim
    assert not bad.any(), f"{what}: {bad.sum()} violations, max err {err.max():.3e}"


@pytest.mark.parametrize("N,M,kw", [(256, 256, 31), (600, 256, 31)])
@pytest.mark.parametrize("prec", ["fp32", "mixed_fp16", "bf16_exchange"])
@pytest.mark.parametrize("rows", [16, 8, 4])
def test_band_heights_and_chunks(orc, N, M, kw, prec, rows):
    e = (kw - 1) // 2
    W = N - 2 * e
    sigma = e / 3.0
    gx, gy, hx, hy = srk.kernel_spectra_srsrk(kw, sigma, M, with_spatial=True)
    Kx, Ky = hx.numpy().astype(np.float64), hy.numpy().astype(np.float64)
    B = 2
    f = patterns.grf(N, 11, B=B)
    g = patterns.grf(N, 12, B=B)
    v = np.random.default_rng(rows).standard_normal((B, W, W)).astype(np.float32)
    ws = srk.Workspace(N, W, e, M, B_max=B, precision=prec)
    ws.set_tuning(chunk_tiles=3, p2_rows=rows)
    h = host(srk.forward_batched(dev(g), dev(f), gx, gy, e, 1.0, ws))
    lf, lg = srk.adjoint_batched(dev(v), dev(g), dev(f), gx, gy, e, 1.0, ws)
    lf, lg = host(lf), host(lg)
    for b in range(B):
        assert_close(h[b], orc.forward(f[b], g[b], e, Kx, Ky, fft=True), prec, f"h[{b}]")
        rf, rg = orc.adjoint(v[b], f[b], g[b], e, Kx, Ky, fft=True)
        assert_close(lf[b], rf, prec, f"lambda_f[{b}]")
        assert_close(lg[b], rg, prec, f"lambda_g[{b}]")


def test_two_stream_adjoint_equals_one_stream():
    """Chunks alternating between two streams (each with its own half of the
    exchange buffers) give bit-identical results to one stream, for the
    forward and the adjoint, eagerly and when the forked call is captured
    into a graph and replayed."""
    import torch
    N, M, kw, B = 600, 256, 31, 3
    e = (kw - 1) // 2
    W = N - 2 * e
    gx, gy = srk.kernel_spectra_srsrk(kw, 5.0, M)
    f = dev(patterns.grf(N, 21, B=B))
    g = dev(patterns.grf(N, 22, B=B))
    v = dev(np.random.default_rng(3).standard_normal((B, W, W)))
    ws = srk.Workspace(N, W, e, M, B_max=B)
    ws.set_tuning(chunk_tiles=1)  # 27 tiles per call in one-tile chunks: the calls take both lanes
    out, hs = {}, {}
    for conc in (0, 1):
        ws.set_tuning(concurrency=conc)
        hs[conc] = srk.forward_batched(g, f, gx, gy, e, 1.0, ws)
        lf, lg = srk.adjoint_batched(v, g, f, gx, gy, e, 1.0, ws)
        torch.cuda.synchronize()
        out[conc] = (lf.clone(), lg.clone())
    assert torch.equal(hs[0], hs[1])
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
    lf = torch.empty_like(out[1][0])
    lg = torch.empty_like(out[1][1])
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            srk.adjoint_batched(v, g, f, gx, gy, e, 1.0, ws, lf, lg)
    torch.cuda.current_stream().wait_stream(s)
    for _ in range(3):
        lf.zero_()
        lg.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(lf, out[0][0]) and torch.equal(lg, out[0][1])


@pytest.mark.parametrize("scratch_mb", [1, 3, 5])
def test_small_exchange_budgets_match_the_default(scratch_mb):
    """Exchange budgets of about one tile: the two-lane split is used only when
    half of the buffers holds a tile, and every chunking gives the same bits
    (tiles are independent)."""
    import torch
    N, M, kw, B = 600, 256, 31, 2
    e = (kw - 1) // 2
    W = N - 2 * e
    gx, gy = srk.kernel_spectra_srsrk(kw, 5.0, M)
    f = dev(patterns.grf(N, 41, B=B))
    g = dev(patterns.grf(N, 42, B=B))
    v = dev(np.random.default_rng(4).standard_normal((B, W, W)))
    res = []
    for mb in (0, scratch_mb):
        ws = srk.Workspace(N, W, e, M, B_max=B, scratch_mb=mb)
        h = srk.forward_batched(g, f, gx, gy, e, 1.0, ws)
        lf, lg = srk.adjoint_batched(v, g, f, gx, gy, e, 1.0, ws)
        torch.cuda.synchronize()
        res.append((h, lf, lg, ws.info()))
    assert res[1][3]["chunk_adj"] <= 2
    for a, b in zip(res[0][:3], res[1][:3]):
        assert torch.equal(a, b)
