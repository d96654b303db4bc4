"""Verification harness with the semantics of the reference's stTestCore
testers, for the operators of this package (or any callables of the same
shape), plus the reference dump formats so that dumps load in the reference
notebooks.

Restated from third_party/include/stTestCore/:
  dump_buffer           dump.h:35-84   (# name / # dtype / # shape W H, then
                                        H lines of W values "%.8g", column-major
                                        source, numpy.loadtxt-compatible)
  dump_adjoint_pixels   dump.h:86-112  (# label, header, "nx ny a n rel abs")
  compare               comparator.h:22-70 (l2, max, argmax position, stdev of
                                        |ref - test|; passed = max <= tolerance)
  adjoint_test          adjoint_tester.h:28-140 (q = sum B^2, analytical
                                        adjoint(2B), central difference at
                                        sampled pixels, mixed tolerance)
  tangent_test          tangent_tester.h:27-62, 447-560 (epsilon sweep, pass if
                                        some epsilon passes every pixel, best
                                        epsilon = fewest violations, then
                                        smallest max abs error)
  tangent_adjoint_test  tangent_adjoint_tester.h:28-128 (same criterion on the
                                        central difference of the adjoint)

Arrays are numpy arrays or torch tensors in the reference's column-major
layout: a 2-D array of shape (W, H) holds column x in row x (element [x, y]
is pixel (nx = x, ny = y)); batches add a leading axis.  Reductions run in
float64 on the host.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, List, Optional, Sequence

import numpy as np

__all__ = [
    "dump_buffer", "load_buffer", "dump_adjoint_pixels", "compare", "CompareResult",
    "AdjointConfig", "AdjointResult", "PixelResult", "adjoint_test",
    "TangentConfig", "TangentResult", "EpsilonResult", "tangent_test", "tangent_adjoint_test",
    "default_epsilon_sweep",
]


def _host(a) -> np.ndarray:
    if hasattr(a, "detach"):
        a = a.detach()
        if a.is_cuda:
            import torch
            torch.cuda.synchronize()
        a = a.cpu().numpy()
    return np.asarray(a)


# ---- dumps ------------------------------------------------------------------

_DTYPES = {"float16": np.float16, "float32": np.float32, "float64": np.float64,
           "int32": np.int32, "int16": np.int16, "int8": np.int8, "uint32": np.uint32,
           "uint16": np.uint16, "uint8": np.uint8}


def dump_buffer(path: str, a, name: str = "", dtype: Optional[str] = None) -> None:
    """Write a W x H buffer (array [W, H], column-major source) as text:
    optional '# name:', '# dtype:', '# shape: W H', then H lines of W values."""
    x = _host(a)
    if x.ndim == 1:
        x = x[:, None]
    if x.ndim != 2:
        raise ValueError("dump_buffer takes one 1-D or 2-D buffer")
    dtype = dtype or str(x.dtype)
    if dtype not in _DTYPES:
        raise ValueError(f"unsupported dtype {dtype}")
    W, H = x.shape
    v = x.astype(np.float64)
    with open(path, "w") as f:
        if name:
            f.write(f"# name: {name}\n")
        f.write(f"# dtype: {dtype}\n")
        f.write(f"# shape: {W} {H}\n")
        for row in range(H):
            f.write(" ".join("%.8g" % v[col, row] for col in range(W)) + "\n")


def load_buffer(path: str):
    """Read a dump_buffer file: (array [W, H] float64, {'name', 'dtype', 'shape'})."""
    meta = {}
    rows = []
    with open(path) as f:
        for line in f:
            if line.startswith("#"):
                k, _, v = line[1:].partition(":")
                meta[k.strip()] = v.strip()
            elif line.strip():
                rows.append([float(t) for t in line.split()])
    W, H = (int(t) for t in meta["shape"].split())
    arr = np.array(rows, dtype=np.float64).reshape(H, W).T
    meta["shape"] = (W, H)
    return arr, meta


@dataclasses.dataclass
class PixelResult:
    nx: int
    ny: int
    analytical: float
    numerical: float
    rel_error: float
    abs_error: float


def dump_adjoint_pixels(path: str, result, label: str = "") -> None:
    """Per-pixel table '# nx ny analytical numerical rel_error abs_error'."""
    with open(path, "w") as f:
        if label:
            f.write(f"# label: {label}\n")
        f.write("# nx ny analytical numerical rel_error abs_error\n")
        for p in result.pixels:
            f.write("%d %d %.8g %.8g %.8g %.8g\n" % (p.nx, p.ny, p.analytical, p.numerical,
                                                      p.rel_error, p.abs_error))


# ---- comparison ---------------------------------------------------------------

@dataclasses.dataclass
class CompareResult:
    l2_error: float
    max_error: float
    max_pos: tuple
    stdev: float
    num_elements: int
    passed: bool
    tolerance: float


def compare(ref, test, tolerance: float = -1.0) -> CompareResult:
    """Elementwise |ref - test| statistics; passed = max <= tolerance (if >= 0).
    The position of the maximum is (nx, ny) of a column-major 2-D buffer."""
    r = _host(ref).astype(np.float64)
    t = _host(test).astype(np.float64)
    d = np.abs(r - t)
    flat = d.reshape(-1)
    i = int(np.argmax(flat)) if flat.size else 0
    if d.ndim >= 2:
        pos = tuple(int(v) for v in np.unravel_index(i, d.shape)[-2:])
    else:
        pos = (i, 0)
    mx = float(flat[i]) if flat.size else 0.0
    return CompareResult(l2_error=float(np.sqrt((flat ** 2).sum())), max_error=mx, max_pos=pos,
                         stdev=float(flat.std()), num_elements=int(flat.size),
                         passed=(mx <= tolerance) if tolerance >= 0 else True,
                         tolerance=tolerance)


def _err(a: float, n: float, tol_abs: float, tol_rel: float):
    ae = abs(a - n)
    scale = max(abs(a), abs(n))
    re = ae / scale if scale > 0 else 0.0
    return ae, re, ae <= max(tol_abs, tol_rel * scale)


def _sample_index(shape, num_samples: int, seed: int):
    """Flat indices (sorted) of the pixels to check in a (W, H) buffer; all
    pixels when num_samples < 0 or >= W*H."""
    W, H = shape[-2], shape[-1]
    total = W * H
    if num_samples < 0 or num_samples >= total:
        return np.arange(total)
    rng = np.random.default_rng(seed if seed else None)
    return np.sort(rng.choice(total, size=num_samples, replace=False))


def _sample(shape, num_samples: int, seed: int):
    H = shape[-1]
    return [(int(i // H), int(i % H)) for i in _sample_index(shape, num_samples, seed)]


# ---- adjoint ------------------------------------------------------------------

@dataclasses.dataclass
class AdjointConfig:
    num_samples: int = 256   # pixels to check; -1 = all
    epsilon: float = 1e-3
    tol_abs: float = 1e-5
    tol_rel: float = 1e-3
    seed: int = 0            # 0 = random


@dataclasses.dataclass
class AdjointResult:
    pixels: List[PixelResult]
    rel_error_mean: float
    rel_error_max: float
    rel_error_stdev: float
    abs_error_mean: float
    abs_error_max: float
    abs_error_stdev: float
    worst_nx: int
    worst_ny: int
    num_samples: int
    passed: bool
    tol_rel: float
    tol_abs: float


def _q(b) -> float:
    """q = sum b^2 in float64 (on the device for device tensors, so large
    outputs are not copied to the host per perturbed pixel)."""
    if hasattr(b, "detach") and b.is_cuda:
        x = b.detach().double()
        return float((x * x).sum())
    x = _host(b).astype(np.float64)
    return float((x * x).sum())


def adjoint_test(forward: Callable, adjoint: Callable, x, cfg: AdjointConfig = AdjointConfig()):
    """Check adjoint against q(x) = sum forward(x)^2: analytical dq/dx =
    adjoint(2 forward(x)), numerical (q(x + eps e_ij) - q(x - eps e_ij)) / (2 eps)
    at cfg.num_samples pixels of the 2-D input x (torch tensor or numpy array;
    forward and adjoint take and return the same kind)."""
    b = forward(x)
    ana = _host(adjoint(2 * b, x)).astype(np.float64)
    pix = _sample(x.shape, cfg.num_samples, cfg.seed)
    out = []
    for nx, ny in pix:
        xp = x.clone() if hasattr(x, "clone") else x.copy()
        xm = x.clone() if hasattr(x, "clone") else x.copy()
        xp[..., nx, ny] += cfg.epsilon
        xm[..., nx, ny] -= cfg.epsilon
        num = (_q(forward(xp)) - _q(forward(xm))) / (2 * cfg.epsilon)
        a = float(ana[..., nx, ny])
        ae, re, _ = _err(a, num, cfg.tol_abs, cfg.tol_rel)
        out.append(PixelResult(nx, ny, a, num, re, ae))
    return _summarise_adjoint(out, cfg)


def _summarise_adjoint(pix, cfg) -> AdjointResult:
    ae = np.array([p.abs_error for p in pix])
    re = np.array([p.rel_error for p in pix])
    ok = [_err(p.analytical, p.numerical, cfg.tol_abs, cfg.tol_rel)[2] for p in pix]
    w = int(np.argmax(re)) if pix else 0
    return AdjointResult(pixels=pix, rel_error_mean=float(re.mean()), rel_error_max=float(re.max()),
                         rel_error_stdev=float(re.std()), abs_error_mean=float(ae.mean()),
                         abs_error_max=float(ae.max()), abs_error_stdev=float(ae.std()),
                         worst_nx=pix[w].nx, worst_ny=pix[w].ny, num_samples=len(pix),
                         passed=all(ok), tol_rel=cfg.tol_rel, tol_abs=cfg.tol_abs)


# ---- tangent / tangent-adjoint ------------------------------------------------

def default_epsilon_sweep() -> List[float]:
    return [1e-4, 2e-4, 5e-4, 1e-3, 2e-3, 5e-3, 1e-2, 2e-2, 5e-2, 1e-1]


@dataclasses.dataclass
class TangentConfig:
    num_samples: int = -1    # pixels to check; -1 = all
    epsilons: Sequence[float] = dataclasses.field(default_factory=default_epsilon_sweep)
    tol_abs: float = 1e-5
    tol_rel: float = 1e-3
    seed: int = 0


@dataclasses.dataclass
class EpsilonResult:
    epsilon: float
    num_violations: int
    abs_error_max: float
    rel_error_max: float
    worst_nx: int
    worst_ny: int
    passed: bool


@dataclasses.dataclass
class TangentResult:
    per_epsilon: List[EpsilonResult]
    passed: bool
    best_epsilon: float
    best_abs_error_max: float
    best_rel_error_max: float
    best_num_violations: int
    num_samples: int
    # per-pixel data at the best epsilon (flat pixel index, analytical,
    # numerical), kept as arrays so that full-image checks of large images
    # stay cheap; `pixels` builds the PixelResult list on demand
    _flat: np.ndarray = dataclasses.field(default=None, repr=False)
    _ana: np.ndarray = dataclasses.field(default=None, repr=False)
    _num: np.ndarray = dataclasses.field(default=None, repr=False)
    _H: int = 1

    @property
    def pixels(self) -> List[PixelResult]:
        out = []
        for i, a, n in zip(self._flat, self._ana, self._num):
            ae = abs(float(a) - float(n))
            sc = max(abs(float(a)), abs(float(n)))
            out.append(PixelResult(int(i // self._H), int(i % self._H), float(a), float(n),
                                   ae / sc if sc > 0 else 0.0, ae))
        return out


def _directional_check(fn: Callable, state: Sequence, dirs: Sequence, analytical,
                       cfg: TangentConfig) -> TangentResult:
    ana = _host(analytical).astype(np.float64)
    H = ana.shape[-1]
    flat = _sample_index(ana.shape, cfg.num_samples, cfg.seed)
    full = flat.size == ana.shape[-2] * H
    a = ana.reshape(-1) if full else ana.reshape(-1)[flat]
    per, nums = [], []
    for eps in cfg.epsilons:
        sp = [s + eps * d if d is not None else s for s, d in zip(state, dirs)]
        sm = [s - eps * d if d is not None else s for s, d in zip(state, dirs)]
        num = (_host(fn(*sp)).astype(np.float64) - _host(fn(*sm)).astype(np.float64)) / (2 * eps)
        n = num.reshape(-1) if full else num.reshape(-1)[flat]
        del num
        ae = np.abs(a - n)
        scale = np.maximum(np.abs(a), np.abs(n))
        re = np.divide(ae, scale, out=np.zeros_like(ae), where=scale > 0)
        bad = ae > np.maximum(cfg.tol_abs, cfg.tol_rel * scale)
        w = int(np.argmax(ae)) if ae.size else 0
        per.append(EpsilonResult(eps, int(bad.sum()), float(ae.max(initial=0.0)),
                                 float(re.max(initial=0.0)), int(flat[w] // H) if ae.size else 0,
                                 int(flat[w] % H) if ae.size else 0, not bad.any()))
        nums.append(n)
    best, any_pass = 0, False
    for k, er in enumerate(per):
        if er.passed:
            if not any_pass or er.abs_error_max < per[best].abs_error_max:
                best, any_pass = k, True
        elif not any_pass:
            if (er.num_violations < per[best].num_violations or
                    (er.num_violations == per[best].num_violations and
                     er.abs_error_max < per[best].abs_error_max)):
                best = k
    b = per[best]
    return TangentResult(per_epsilon=per, passed=any_pass, best_epsilon=b.epsilon,
                         best_abs_error_max=b.abs_error_max, best_rel_error_max=b.rel_error_max,
                         best_num_violations=b.num_violations, num_samples=int(flat.size),
                         _flat=flat, _ana=a, _num=nums[best], _H=H)


def tangent_test(forward: Callable, tangent: Callable, state: Sequence, dirs: Sequence,
                 cfg: TangentConfig = TangentConfig()) -> TangentResult:
    """Analytical tangent(state, dirs) against the central difference of
    forward(*state) along dirs (None = that input held fixed), over an epsilon
    sweep; passes if some epsilon passes every sampled pixel."""
    return _directional_check(forward, state, dirs, tangent(state, dirs), cfg)


def tangent_adjoint_test(adjoint: Callable, tangent_adjoint: Callable, state: Sequence,
                         dirs: Sequence, cfg: TangentConfig = TangentConfig()) -> TangentResult:
    """Analytical tangent_adjoint(state, dirs) against the central difference of
    adjoint(*state) (one adjoint output) along dirs, as tangent_test."""
    return _directional_check(adjoint, state, dirs, tangent_adjoint(state, dirs), cfg)
