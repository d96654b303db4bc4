"""Synthetic test images (host side), column-major: arrays [.., N, N] indexed
[.., column, row].

- ``lns_pulse_train``: the reference's line/space pulse train
  (third_party/src/patternUtil/test_pattern.cu:67-109): a sum of Gaussian
  pulses of width sigma at pitch ``period`` (default 4*FWHM) around the image
  centre, minus 0.5, constant along the lines.
- ``grf``: Gaussian random field -- seeded N(0,1) white noise, periodic
  Gaussian smoothing of width sigma, standardised, then sigmoid to (0, 1).
- ``lns_box``: line/space of pitch 64 px, duty 0.5, box edges blurred by a
  Gaussian of sigma 2 px, values in [0, 1].
- ``gp_field`` / ``gp_fields``: the reference's Gaussian-process fixtures
  (GpGenerator, third_party/include/stTestCore/gp_generator.h:95-201), e.g.
  ``gp_field(N, 32.0, 42)`` / ``(N, 32.0, 43)`` for the f / g inputs of
  tan_tanadj/main_tan_tanadj.cu:322-323; generated by srkGpGenerate in the
  native library (host code, same mt19937_64 / normal_distribution stream).
"""
from __future__ import annotations

import math

import numpy as np


def lns_pulse_train(N: int, sigma: float, period: float = 0.0, nbrs: int = 4,
                    vertical: bool = True) -> np.ndarray:
    fwhm = 2.0 * math.sqrt(2.0 * math.log(2.0)) * sigma
    p = period if period > 0 else 4.0 * fwhm
    centre = 0.5 * (N - 1)
    x = np.arange(N, dtype=np.float64)
    val = np.zeros(N)
    for k in range(-nbrs, nbrs + 1):
        d = x - (centre + k * p)
        val += np.exp(-d * d / (2.0 * sigma * sigma))
    val -= 0.5
    img = np.empty((N, N), np.float64)
    if vertical:      # varies with the column index
        img[:, :] = val[:, None]
    else:             # varies with the row index
        img[:, :] = val[None, :]
    return img.astype(np.float32)


def _gauss_filter(N: int, sigma: float) -> np.ndarray:
    k = np.fft.fftfreq(N)
    return np.exp(-2.0 * math.pi ** 2 * sigma ** 2 * k * k)


def grf(N: int, seed: int, sigma: float = 8.0, B: int = 1) -> np.ndarray:
    """[B, N, N] float32 Gaussian random fields in (0, 1)."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((B, N, N))
    g1 = _gauss_filter(N, sigma)
    H = g1[:, None] * g1[None, :]
    zs = np.fft.ifft2(np.fft.fft2(z) * H).real
    zs -= zs.mean(axis=(1, 2), keepdims=True)
    zs /= zs.std(axis=(1, 2), keepdims=True)
    return (1.0 / (1.0 + np.exp(-zs))).astype(np.float32)


def lns_box(N: int, pitch: float = 64.0, duty: float = 0.5, sigma: float = 2.0,
            vertical: bool = True) -> np.ndarray:
    """[N, N] float32 blurred box line/space pattern, lines centred on the image centre."""
    x = np.arange(N, dtype=np.float64)
    centre = 0.5 * (N - 1)
    ph = np.mod(x - centre + 0.5 * pitch, pitch) - 0.5 * pitch
    box = (np.abs(ph) < 0.5 * duty * pitch).astype(np.float64)
    prof = np.fft.ifft(np.fft.fft(box) * _gauss_filter(N, sigma)).real
    img = np.empty((N, N), np.float64)
    if vertical:
        img[:, :] = prof[:, None]
    else:
        img[:, :] = prof[None, :]
    return img.astype(np.float32)


def grf_torch(B: int, N: int, seed: int, sigma: float = 8.0, device="cuda"):
    """GPU version of ``grf`` (same recipe, torch RNG) for large batches."""
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    k = torch.fft.fftfreq(N, device=device, dtype=torch.float64)
    g1 = torch.exp(-2.0 * math.pi ** 2 * sigma ** 2 * k * k).float()
    H = (g1[:, None] * g1[None, :]).to(torch.complex64)
    out = torch.empty((B, N, N), dtype=torch.float32, device=device)
    step = max(1, (1 << 26) // (N * N))
    for b0 in range(0, B, step):
        b1 = min(B, b0 + step)
        z = torch.randn((b1 - b0, N, N), generator=gen, device=device)
        zs = torch.fft.ifft2(torch.fft.fft2(z) * H).real
        zs = zs - zs.mean(dim=(1, 2), keepdim=True)
        zs = zs / zs.std(dim=(1, 2), keepdim=True)
        out[b0:b1] = torch.sigmoid(zs)
    return out


def gp_fields(W: int, H: int, ell_x: float, seed: int, count: int = 1, ell_y: float = 0.0,
              rho: float = 0.0, sigma: float = 1.0) -> np.ndarray:
    """[count, W, H] float32 consecutive GP fields of one seeded stream
    (column-major: [x, y] indexing), covariance sigma^2 exp(-Q/2)."""
    import ctypes
    from ._lib import call
    out = np.empty((count, W, H), np.float32)
    call("srkGpGenerate", float(sigma), float(ell_x), float(ell_y or ell_x), float(rho), W, H,
         ctypes.c_ulonglong(seed), count, out.ctypes.data)
    return out


def gp_field(N: int, length_scale_px: float = 32.0, seed: int = 42) -> np.ndarray:
    """TestPatternConfig::gaussian_process(N, length_scale_px, seed): [N, N] float32."""
    return gp_fields(N, N, length_scale_px, seed)[0]
