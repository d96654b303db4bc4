"""FP64 host oracle for the SRK shrinkage operator -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker (or the timed CPU
baseline); the product package never imports it.

``srk_oracle.c`` holds the arithmetic (double precision, pthreads); ``cd.py``
holds the host restatement of the reference CD metric.  Every function cites
the reference file:line it restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsrk_oracle.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)


def build(force: bool = False) -> str:
    """Compile ``libsrk_oracle.so`` with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_bessel_k1.restype = ctypes.c_double
        L.orc_bessel_k1.argtypes = [ctypes.c_double]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_make_srsrk_kernel.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, _D, _D]
        i, d = ctypes.c_int, ctypes.c_double
        L.orc_conv_valid.argtypes = [_D, i, i, _D, _D, i, d, _D, _D]
        L.orc_forward.argtypes = [_D, _D, i, i, _D, _D, i, d, _D]
        L.orc_forward_fft.argtypes = [_D, _D, i, i, _D, _D, i, d, _D]
        L.orc_adjoint_f_from_u.argtypes = [_D, _D, _D, i, i, _D]
        L.orc_adjoint_g.argtypes = [_D, _D, i, i, _D, _D, i, d, _D]
        L.orc_adjoint.argtypes = [_D, _D, _D, i, i, _D, _D, i, d, _D, _D]
        L.orc_adjoint_fft.argtypes = [_D, _D, _D, i, i, _D, _D, i, d, _D, _D]
        L.orc_tangent.argtypes = [_D, _D, _D, _D, i, i, _D, _D, i, d, _D]
        L.orc_tangent_adjoint.argtypes = [_D, _D, _D, _D, _D, _D, i, i, _D, _D, i, d,
                                          _D, _D]
        L.orc_forward_window.argtypes = [_D, _D, i, i, _D, _D, i, d, i, i, i, i, _D]
        L.orc_adjoint_window.argtypes = [_D, _D, _D, i, i, _D, _D, i, d, i, i, i, i, _D, _D]
        L.orc_fwd_adj_fft_batch.argtypes = [_D, _D, _D, i, i, i, _D, _D, i, d, _D, _D, _D]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def threads() -> int:
    return int(lib().orc_get_threads())


# Arrays below are numpy (C-order) views of column-major images: a[c, r] is
# the pixel at column c (x) and row r (y), i.e. flat index c*H + r.

def bessel_k1(x: float) -> float:
    return float(lib().orc_bessel_k1(float(x)))


def srsrk_kernel(kw: int, dx: float = 1.0, sigma: float = 5.0, cw: float = 3.0):
    """(Kx, Ky) as [KW, KW] arrays indexed [i (x offset), j (y offset)]."""
    Kx = np.zeros((kw, kw), np.float64)
    Ky = np.zeros((kw, kw), np.float64)
    lib().orc_make_srsrk_kernel(int(kw), float(dx), float(sigma), float(cw), _p(Kx), _p(Ky))
    return Kx, Ky


def conv_valid(g, e, Kx, Ky, s=1.0):
    g = _f64(g); Kx = _f64(Kx); Ky = _f64(Ky)
    N = g.shape[0]; W = N - 2 * e
    ux = np.zeros((W, W)); uy = np.zeros((W, W))
    lib().orc_conv_valid(_p(g), N, int(e), _p(Kx), _p(Ky), Kx.shape[0], float(s), _p(ux), _p(uy))
    return ux, uy


def forward(f, g, e, Kx, Ky, s=1.0, fft=False):
    f = _f64(f); g = _f64(g); Kx = _f64(Kx); Ky = _f64(Ky)
    N = f.shape[0]; W = N - 2 * e
    h = np.zeros((W, W))
    fn = lib().orc_forward_fft if fft else lib().orc_forward
    fn(_p(f), _p(g), N, int(e), _p(Kx), _p(Ky), Kx.shape[0], float(s), _p(h))
    return h


def adjoint_f_from_u(v, ux, uy, N, e):
    v = _f64(v); ux = _f64(ux); uy = _f64(uy)
    lam = np.zeros((N, N))
    lib().orc_adjoint_f_from_u(_p(v), _p(ux), _p(uy), int(N), int(e), _p(lam))
    return lam


def adjoint_g(v, f, e, Kx, Ky, s=1.0):
    v = _f64(v); f = _f64(f); Kx = _f64(Kx); Ky = _f64(Ky)
    N = f.shape[0]
    lam = np.zeros((N, N))
    lib().orc_adjoint_g(_p(v), _p(f), N, int(e), _p(Kx), _p(Ky), Kx.shape[0], float(s), _p(lam))
    return lam


def adjoint(v, f, g, e, Kx, Ky, s=1.0, fft=False):
    v = _f64(v); f = _f64(f); g = _f64(g); Kx = _f64(Kx); Ky = _f64(Ky)
    N = f.shape[0]
    lf = np.zeros((N, N)); lg = np.zeros((N, N))
    fn = lib().orc_adjoint_fft if fft else lib().orc_adjoint
    fn(_p(v), _p(f), _p(g), N, int(e), _p(Kx), _p(Ky), Kx.shape[0], float(s), _p(lf), _p(lg))
    return lf, lg


def tangent(f, g, df, dg, e, Kx, Ky, s=1.0):
    f, g, df, dg = (_f64(a) for a in (f, g, df, dg))
    Kx = _f64(Kx); Ky = _f64(Ky)
    N = f.shape[0]; W = N - 2 * e
    dh = np.zeros((W, W))
    lib().orc_tangent(_p(f), _p(g), _p(df), _p(dg), N, int(e), _p(Kx), _p(Ky), Kx.shape[0],
                      float(s), _p(dh))
    return dh


def tangent_adjoint(f, g, df, dg, lam_h, dlam_h, e, Kx, Ky, s=1.0):
    f, g, df, dg, lam_h, dlam_h = (_f64(a) for a in (f, g, df, dg, lam_h, dlam_h))
    Kx = _f64(Kx); Ky = _f64(Ky)
    N = f.shape[0]
    a = np.zeros((N, N)); b = np.zeros((N, N))
    lib().orc_tangent_adjoint(_p(f), _p(g), _p(df), _p(dg), _p(lam_h), _p(dlam_h), N, int(e),
                              _p(Kx), _p(Ky), Kx.shape[0], float(s), _p(a), _p(b))
    return a, b


def fwd_adj_fft_batch(f, g, v, e, Kx, Ky, s=1.0):
    """FP64 FFT-form forward + adjoint over a batch [B, N, N] (CPU baseline)."""
    f = _f64(f); g = _f64(g); v = _f64(v); Kx = _f64(Kx); Ky = _f64(Ky)
    B, N = f.shape[0], f.shape[1]; W = N - 2 * e
    h = np.zeros((B, W, W)); lf = np.zeros((B, N, N)); lg = np.zeros((B, N, N))
    lib().orc_fwd_adj_fft_batch(_p(f), _p(g), _p(v), B, N, int(e), _p(Kx), _p(Ky), Kx.shape[0],
                                float(s), _p(h), _p(lf), _p(lg))
    return h, lf, lg


def forward_window(f, g, e, Kx, Ky, ox0, oy0, wx, wy, s=1.0):
    """h on output window [ox0, ox0+wx) x [oy0, oy0+wy) (valid-region coordinates)."""
    f = _f64(f); g = _f64(g); Kx = _f64(Kx); Ky = _f64(Ky)
    out = np.zeros((wx, wy))
    lib().orc_forward_window(_p(f), _p(g), f.shape[0], int(e), _p(Kx), _p(Ky), Kx.shape[0],
                             float(s), int(ox0), int(oy0), int(wx), int(wy), _p(out))
    return out


def adjoint_window(v, f, g, e, Kx, Ky, qx0, qy0, wx, wy, s=1.0):
    """(lambda_f, lambda_g) on image window [qx0, qx0+wx) x [qy0, qy0+wy)."""
    v = _f64(v); f = _f64(f); g = _f64(g); Kx = _f64(Kx); Ky = _f64(Ky)
    a = np.zeros((wx, wy)); b = np.zeros((wx, wy))
    lib().orc_adjoint_window(_p(v), _p(f), _p(g), f.shape[0], int(e), _p(Kx), _p(Ky), Kx.shape[0],
                             float(s), int(qx0), int(qy0), int(wx), int(wy), _p(a), _p(b))
    return a, b


def tangent_window(f, g, df, dg, e, Kx, Ky, ox0, oy0, wx, wy, s=1.0):
    """Tangent dh on an output window (valid-region coordinates).

    The forward is bilinear in (f, g), so its tangent
    dh = u_x(dg) f_x + u_y(dg) f_y + u_x(g) df_x + u_y(g) df_y
    (tan_tanadj/srk_tangent.cuh:28-52) is h(f, dg) + h(df, g), evaluated here
    with two direct forward windows."""
    return (forward_window(f, dg, e, Kx, Ky, ox0, oy0, wx, wy, s) +
            forward_window(df, g, e, Kx, Ky, ox0, oy0, wx, wy, s))


def tangent_adjoint_window(f, g, df, dg, lam_h, dlam_h, e, Kx, Ky, qx0, qy0, wx, wy, s=1.0):
    """(dLam_f, dLam_g) on an image window.

    dLam_f = A_f(dlam_h; u(g)) + A_f(lam_h; u(dg)) and
    dLam_g = A_g(dlam_h; f) + A_g(lam_h; df)
    (tan_tanadj/srk_tangent_adjoint.cuh:100-159), i.e. the adjoint window of
    (dlam_h, f, g) plus the adjoint window of (lam_h, df, dg)."""
    a1, b1 = adjoint_window(dlam_h, f, g, e, Kx, Ky, qx0, qy0, wx, wy, s)
    a2, b2 = adjoint_window(lam_h, df, dg, e, Kx, Ky, qx0, qy0, wx, wy, s)
    return a1 + a2, b1 + b2
