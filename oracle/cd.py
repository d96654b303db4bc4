"""Host restatement of the reference CD metric -- TEST INFRASTRUCTURE ONLY.

Follows third_party/src/metricUtil/cd_measure_util.cu:
  - sinc x Blackman polyphase table, L phases, 2A+1 taps, rows normalised to
    sum 1 (:43-87);
  - upsample out[k*L + r] = sum_i sig[k + i - A] h[r][i] (:96-125);
  - zero crossings between consecutive samples, linear interpolation
    x = i - y0/(y1 - y0) (:134-170);
  - nearest crossing left of / at-or-right of the centre (N*L - 1)/2 (:205-256,
    centre at :421);
  - CD = (right - left)/L pixels (:264-282);
  - a missing edge keeps the search's start value -/+FLT_MAX (:37, 218-219), so
    it reads -/+FLT_MAX/L pixels, and cd = right - left in FP32 (:264-282).
Found edges are computed in double precision; the missing-edge values follow
the reference's FP32 arithmetic exactly.
"""
from __future__ import annotations

import numpy as np


def _sinc_blackman(t: np.ndarray, A: int) -> np.ndarray:
    at = np.abs(t)
    s = np.where(at < 1e-6, 1.0, np.sin(np.pi * t) / np.where(at < 1e-6, 1.0, np.pi * t))
    u = (t + A) / (2.0 * A)
    bw = 0.42 - 0.5 * np.cos(2 * np.pi * u) + 0.08 * np.cos(4 * np.pi * u)
    return np.where(at > A + 1e-6, 0.0, s * bw)


def polyphase_table(L: int = 16, A: int = 8) -> np.ndarray:
    n = np.arange(-A, A + 1, dtype=np.float64)
    ph = np.arange(L, dtype=np.float64) / L
    h = _sinc_blackman(n[None, :] - ph[:, None], A)
    s = h.sum(axis=1, keepdims=True)
    return h / np.where(np.abs(s) > 1e-20, s, 1.0)


def upsample(sig: np.ndarray, L: int = 16, A: int = 8) -> np.ndarray:
    """sig [B, N] -> [B, N*L]."""
    sig = np.asarray(sig, np.float64)
    B, N = sig.shape
    h = polyphase_table(L, A)
    pad = np.zeros((B, N + 2 * A))
    pad[:, A:A + N] = sig
    out = np.zeros((B, N, L))
    for i in range(2 * A + 1):
        # sig[k + i - A] == pad[k + i]
        out += pad[:, i:i + N, None] * h[None, None, :, i]
    return out.reshape(B, N * L)


def measure_cd(sig: np.ndarray, L: int = 16, A: int = 8, threshold: float = 0.0):
    """CD in pixels of each row of sig [B, N] at the given threshold.
    Returns (left_px, right_px, cd_px); a missing crossing gives -/+FLT_MAX/L."""
    up = upsample(np.asarray(sig, np.float64) - threshold, L, A)
    B, Nup = up.shape
    y0, y1 = up[:, :-1], up[:, 1:]
    cross = ((y0 <= 0) & (y1 > 0)) | ((y0 > 0) & (y1 <= 0))
    den = y1 - y0
    idx = np.arange(Nup - 1, dtype=np.float64)[None, :]
    x = np.where(np.abs(den) < 1e-30, idx, idx - y0 / np.where(den == 0, 1.0, den))
    centre = 0.5 * (Nup - 1)
    left = np.where(cross & (x < centre), x, -np.inf).max(axis=1)
    right = np.where(cross & (x >= centre), x, np.inf).min(axis=1)
    fl, fr = np.isfinite(left), np.isfinite(right)
    inv = np.float32(1.0) / np.float32(L)
    lpx = np.where(fl, left / L, np.float64(np.float32(-FLT_MAX) * inv))
    rpx = np.where(fr, right / L, np.float64(np.float32(FLT_MAX) * inv))
    cd32 = (rpx.astype(np.float32) - lpx.astype(np.float32)).astype(np.float64)
    cd = np.where(fl & fr, (right - left) / L, cd32)
    return lpx, rpx, cd


FLT_MAX = float(np.finfo(np.float32).max)


def found(px, L: int = 16) -> np.ndarray:
    """True where an edge position (pixels) was found rather than missing."""
    return np.abs(np.asarray(px, np.float64)) < 0.5 * FLT_MAX / L
