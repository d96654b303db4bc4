"""Autotuner for the tile engine: picks the overlap-and-save tile (FFT) size M,
the row-pass band height RY (which also fixes the radix plan of the tile
transforms, one per M), the launch chunk and the two-lane pipelining for a
problem shape by timing the candidates on the device, and caches the choice.

The reference picks its tile size at compile time (the template parameter M
of srkForwardOaS / adjointSRK, srk_forward.cuh:468-501); here the engine takes
M at run time, so the choice can be measured per shape:

    from shrinkage_full_tests_amd import autotune
    best = autotune.autotune(N=4096, kw=61, B=1)          # times the candidates
    ws = autotune.workspace(best, B_max=1)                # tuned workspace
    gx, gy = srk.kernel_spectra_srsrk(61, 10.0, best["M"])

Results are cached per (device, N, KW, B, operators, precision) in a JSON file
(default ~/.cache/srk_amd_autotune.json, or SRK_AUTOTUNE_CACHE).
"""
from __future__ import annotations

import itertools
import json
import os
from typing import Dict, Iterable, List, Optional

import shrinkage_full_tests_amd as srk
from shrinkage_full_tests_amd import patterns

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "srk_amd_autotune.json")


def candidate_tiles(N: int, kw: int) -> List[int]:
    """Tile sizes the engine accepts for N and KW: a whole-image tile (M >= N)
    or overlap-and-save tiles with room for the halo (M > 2e + 2).  2048-point
    overlap-and-save tiles are left out where 1024-point ones fit the halo:
    their passes hold about 70 KB of LDS per block and run slower at every
    size measured (N = 16384: 14.0 against 9.6 ms; N = 32768: 53.9 against
    37.7 ms)."""
    e = (kw - 1) // 2
    out = []
    for M in srk.SUPPORTED_M:
        if M >= N or M > 2 * e + 2:
            out.append(M)
    # a whole-image tile larger than the smallest one that covers N only pads
    whole = [M for M in out if M >= N]
    oas = [M for M in out if M < N]
    if 2048 in oas and 1024 in oas:
        oas.remove(2048)
    return oas + whole[:1]


def candidates(N: int, kw: int, precision: str = "fp32") -> List[Dict]:
    out = []
    for M in candidate_tiles(N, kw):
        if precision != "fp32" and M > 1024:
            continue
        rows = (4, 8, 16) if M <= 512 else (0,)
        for ry, conc in itertools.product(rows, (1, 0)):
            out.append({"M": M, "p2_rows": ry, "concurrency": conc})
    return out


def _key(N, kw, B, ops, precision):
    import torch
    dev = torch.cuda.get_device_name(0)
    return f"{dev}|N={N}|kw={kw}|B={B}|ops={'+'.join(ops)}|{precision}"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def workspace(cfg: Dict, B_max: int = 1) -> "srk.Workspace":
    """A workspace built and tuned with an autotune() result."""
    ws = srk.Workspace(cfg["N"], cfg["W"], cfg["erosion"], cfg["M"], B_max=B_max,
                       precision=cfg.get("precision", "fp32"))
    ws.set_tuning(p2_rows=cfg["p2_rows"], concurrency=cfg["concurrency"])
    return ws


def autotune(N: int, kw: int, B: int = 1, sigma: Optional[float] = None, dx: float = 1.0,
             ops: Iterable[str] = ("forward", "adjoint"), precision: str = "fp32",
             cands: Optional[List[Dict]] = None, reps: int = 3, cache_path: Optional[str] = None,
             use_cache: bool = True) -> Dict:
    """Time every candidate on synthetic inputs of the given shape (seeded
    Gaussian random fields, random cotangent) and return the fastest:
    {"M", "p2_rows", "concurrency", "N", "W", "erosion", "ms", "table"}."""
    import torch
    ops = tuple(ops)
    e = (kw - 1) // 2
    W = N - 2 * e
    sigma = sigma if sigma is not None else e / 3.0
    cache_path = cache_path or os.environ.get("SRK_AUTOTUNE_CACHE", DEFAULT_CACHE)
    key = _key(N, kw, B, ops, precision)
    if use_cache:
        hit = _load(cache_path).get(key)
        if hit:
            return hit
    f = patterns.grf_torch(B, N, seed=11)
    g = patterns.grf_torch(B, N, seed=12)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(13)
    v = torch.randn((B, W, W), generator=gen, device="cuda")
    h = torch.empty((B, W, W), device="cuda")
    lf = torch.empty((B, N, N), device="cuda")
    lg = torch.empty((B, N, N), device="cuda")
    table = []
    spectra = {}
    for c in (cands or candidates(N, kw, precision)):
        M = c["M"]
        if M not in spectra:
            spectra[M] = srk.kernel_spectra_srsrk(kw, sigma, M, dx=dx)
        gx, gy = spectra[M]
        try:
            ws = srk.Workspace(N, W, e, M, B_max=B, precision=precision)
            ws.set_tuning(p2_rows=c["p2_rows"], concurrency=c["concurrency"])
        except srk.SrkError:
            continue

        def step():
            if "forward" in ops:
                srk.forward_batched(g, f, gx, gy, e, dx, ws, h)
            if "adjoint" in ops:
                srk.adjoint_batched(v, g, f, gx, gy, e, dx, ws, lf, lg)

        step()
        torch.cuda.synchronize()
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(reps):
            step()
        t1.record()
        torch.cuda.synchronize()
        table.append(dict(c, ms=t0.elapsed_time(t1) / reps))
        ws.close()
    if not table:
        raise srk.SrkError(-5, "autotune: no candidate tile size fits this shape")
    best = min(table, key=lambda r: r["ms"])
    out = dict(best, N=N, W=W, erosion=e, kw=kw, B=B, precision=precision, table=table)
    if use_cache:
        data = _load(cache_path)
        data[key] = out
        os.makedirs(os.path.dirname(cache_path) or ".", exist_ok=True)
        with open(cache_path, "w") as fh:
            json.dump(data, fh, indent=1)
    return out
