"""The autotuner times the engine's candidates for a shape, caches the
choice, and the tuned workspace computes the same operator (FP64 oracle at
the fp32 tolerance, README.md:8-10)."""
import json

import numpy as np
import pytest

import shrinkage_full_tests_amd as srk
from shrinkage_full_tests_amd import autotune, patterns

pytestmark = pytest.mark.gpu


def test_candidate_tiles_respect_the_halo():
    assert autotune.candidate_tiles(512, 121) == [256, 512]      # 256 > 2*60 + 2
    assert autotune.candidate_tiles(200, 31) == [256]            # whole image only
    # 2048-point overlap-and-save tiles only where 1024-point ones cannot hold the halo
    assert autotune.candidate_tiles(4096, 121) == [256, 512, 1024]
    assert autotune.candidate_tiles(4096, 1023) == [2048]          # 1024 <= 2*511 + 2
    assert autotune.candidate_tiles(2000, 121) == [256, 512, 1024, 2048]  # 2048: whole image


def test_autotune_picks_a_working_configuration(orc, tmp_path):
    import torch
    N, kw, B = 600, 31, 2
    e = (kw - 1) // 2
    W = N - 2 * e
    cache = str(tmp_path / "tune.json")
    cands = [{"M": 256, "p2_rows": 8, "concurrency": 1}, {"M": 256, "p2_rows": 4, "concurrency": 0},
             {"M": 1024, "p2_rows": 0, "concurrency": 1}]
    best = autotune.autotune(N, kw, B, sigma=5.0, cands=cands, reps=1, cache_path=cache)
    assert best["M"] in (256, 1024) and len(best["table"]) == 3
    assert best["ms"] == min(r["ms"] for r in best["table"])
    assert json.load(open(cache))                                  # cached
    again = autotune.autotune(N, kw, B, sigma=5.0, cands=cands[:1], reps=1, cache_path=cache)
    assert again["M"] == best["M"] and len(again["table"]) == 3    # served from the cache
    ws = autotune.workspace(best, B_max=B)
    gx, gy, hx, hy = srk.kernel_spectra_srsrk(kw, 5.0, best["M"], with_spatial=True)
    f = patterns.grf(N, 31, B=B)
    g = patterns.grf(N, 32, B=B)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    h = srk.forward_batched(dev(g), dev(f), gx, gy, e, 1.0, ws).cpu().numpy().astype(np.float64)
    Kx, Ky = hx.numpy().astype(np.float64), hy.numpy().astype(np.float64)
    for b in range(B):
        ref = orc.forward(f[b], g[b], e, Kx, Ky, fft=True)
        lim = np.maximum(1e-5, 1e-3 * np.maximum(np.abs(h[b]), np.abs(ref)))
        assert (np.abs(h[b] - ref) <= lim).all()
