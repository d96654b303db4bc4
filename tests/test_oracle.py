"""Pins the FP64 oracle (CPU only).

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index, a transposed operand) fails at least one of them:
library golden values (scipy), the reference notebook's recorded outputs,
independent library convolutions, adjoint identities and exact central
differences (the operator is bilinear).
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.signal

from shrinkage_full_tests_amd import patterns

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "k1_scipy.json")


def _kernel(orc, kw):
    sigma = (kw // 2) / 3.0   # Rw = 3 sigma = KW//2 (reference default tie, main_cd_training.cu:65,301)
    return orc.srsrk_kernel(kw, 1.0, sigma)


# ---- kernel generation -------------------------------------------------------

def test_k1_matches_scipy(orc):
    g = json.load(open(GOLDEN))
    for x, ref in g["k1"]:
        assert abs(orc.bessel_k1(x) - ref) <= 1e-13 * ref, x


def test_srsrk_samples_match_scipy_built_kernel(orc):
    g = json.load(open(GOLDEN))
    for key, rows in g["srsrk_samples"].items():
        kw, sigma = key.split("_")
        Kx, Ky = orc.srsrk_kernel(int(kw), 1.0, float(sigma))
        for i, j, kx, ky in rows:
            assert abs(Kx[i, j] - kx) <= 1e-13 * max(1e-30, abs(kx)) + 1e-300
            assert abs(Ky[i, j] - ky) <= 1e-13 * max(1e-30, abs(ky)) + 1e-300


def test_kernel_known_answer_notebook(orc):
    # notebooks/plot_batch_demo.ipynb:66-88: sigma=20, KW=31, "Kx peak (r=1 pixel): 0.1583"
    Kx, Ky = orc.srsrk_kernel(31, 1.0, 20.0)
    assert abs(Kx[16, 15] - 0.1583) < 5e-5
    assert abs(Kx.sum()) < 1e-14


def test_kernel_symmetries(orc):
    for kw in (31, 61, 121):
        Kx, Ky = _kernel(orc, kw)
        np.testing.assert_allclose(Kx[::-1, ::-1], -Kx, atol=1e-18)   # odd
        np.testing.assert_allclose(Ky, Kx.T, atol=1e-18)                # Ky = Kx^T
        c = kw // 2
        assert Kx[c, c] == 0.0 and Kx[c + 1, c] > 0 and Ky[c, c + 1] > 0
        # support: zero beyond the Blackman radius Rw = c
        yy, xx = np.meshgrid(np.arange(kw) - c, np.arange(kw) - c)
        assert np.all(Kx[np.hypot(xx, yy) > c] == 0)


# ---- forward -----------------------------------------------------------------

def _rand_fields(N, seed, n=2):
    rng = np.random.default_rng(seed)
    return [rng.random((N, N)) for _ in range(n)]


def _forward_scipy(f, g, Kx, Ky, e, s=1.0):
    """Independent forward: scipy full convolution + numpy central differences."""
    N, kw = f.shape[0], Kx.shape[0]
    c, W = kw // 2, f.shape[0] - 2 * e
    ux = scipy.signal.convolve2d(g, Kx, mode="full")[e + c:e + c + W, e + c:e + c + W] * s
    uy = scipy.signal.convolve2d(g, Ky, mode="full")[e + c:e + c + W, e + c:e + c + W] * s
    fx = np.gradient(f, axis=0)[e:N - e, e:N - e]   # axis 0 = column index = x
    fy = np.gradient(f, axis=1)[e:N - e, e:N - e]
    return ux * fx + uy * fy


@pytest.mark.parametrize("N,kw,e", [(40, 11, 5), (37, 9, 6), (48, 15, 7)])
def test_forward_direct_matches_scipy(orc, N, kw, e):
    Kx, Ky = _kernel(orc, kw)
    f, g = _rand_fields(N, N + kw)
    ref = _forward_scipy(f, g, Kx, Ky, e, s=2.0)
    np.testing.assert_allclose(orc.forward(f, g, e, Kx, Ky, s=2.0), ref, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("N,kw,e", [(40, 11, 5), (48, 15, 8), (64, 31, 16)])
def test_forward_fft_matches_direct(orc, N, kw, e):
    Kx, Ky = _kernel(orc, kw)
    f, g = _rand_fields(N, 7)
    a = orc.forward(f, g, e, Kx, Ky)
    b = orc.forward(f, g, e, Kx, Ky, fft=True)
    np.testing.assert_allclose(b, a, rtol=1e-10, atol=1e-13)


def _half_ulp4(x):
    return 0.5e-3 * 10 ** math.floor(math.log10(abs(x))) * 1.0001


def test_batch_demo_known_answer(orc):
    """training/main_batch_demo.cu with run_batch_demo.sh defaults (sigma=20 ->
    KW=121, erosion=61, N=256, W=134, E=I=LNS pulse train); recorded stats in
    notebooks/plot_batch_demo.ipynb:225-228."""
    N, e = 256, 61
    Kx, Ky = orc.srsrk_kernel(121, 1.0, 20.0)
    expect = {8.0: (-0.234, -0.000, -6.484e-2, 8.204e-2), 6.0: (-0.288, -0.000, -1.029e-1, 1.074e-1)}
    for sig, (mn, mx, mean, std) in expect.items():
        for vertical in (True, False):
            E = patterns.lns_pulse_train(N, sig, vertical=vertical).astype(np.float64)
            R = orc.forward(E, E, e, Kx, Ky, fft=True)
            assert R.shape == (134, 134)
            assert abs(R.min() - mn) < 5e-4 and abs(R.max() - mx) < 5e-4
            # the notebook prints 4 significant digits: allow half a unit of the last one
            assert abs(R.mean() - mean) <= _half_ulp4(mean)
            assert abs(R.std() - std) <= _half_ulp4(std)
            assert R.max() <= 1e-9   # true convolution of odd kernels gives R <= 0 on LNS
    # one direct (non-FFT) evaluation of the same case
    E = patterns.lns_pulse_train(N, 8.0).astype(np.float64)
    R = orc.forward(E, E, e, Kx, Ky)
    assert abs(R.mean() - (-6.484e-2)) <= _half_ulp4(6.484e-2)


# ---- adjoints ----------------------------------------------------------------

@pytest.mark.parametrize("N,kw,e", [(36, 11, 5), (41, 9, 6)])
def test_adjoint_dot_products(orc, N, kw, e):
    """<A_g df, v> = <df, lambda_f> and <B_f dg, v> = <dg, lambda_g> (SRSRK, odd kernels)."""
    Kx, Ky = _kernel(orc, kw)
    f, g, df, dg = _rand_fields(N, 11, 4)
    W = N - 2 * e
    v = np.random.default_rng(3).standard_normal((W, W))
    lam_f, lam_g = orc.adjoint(v, f, g, e, Kx, Ky, s=1.5)
    lhs_f = np.sum(orc.forward(df, g, e, Kx, Ky, s=1.5) * v)
    lhs_g = np.sum(orc.forward(f, dg, e, Kx, Ky, s=1.5) * v)
    assert abs(lhs_f - np.sum(df * lam_f)) <= 1e-11 * abs(lhs_f)
    assert abs(lhs_g - np.sum(dg * lam_g)) <= 1e-11 * abs(lhs_g)
    lf2, lg2 = orc.adjoint(v, f, g, e, Kx, Ky, s=1.5, fft=True)
    np.testing.assert_allclose(lf2, lam_f, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(lg2, lam_g, rtol=1e-9, atol=1e-12)


def test_adjoint_g_sign_against_reference_gpu_form(orc):
    """The reference GPU adjointWrtE computes sum_d conv(v f_d, G_d) (Ghat not
    conjugated, adjoint/srk_adjoint.cuh:84-163).  That equals the true adjoint
    for even kernels and its negative for the odd SRSRK kernels."""
    N, e, kw = 34, 6, 9
    f, = _rand_fields(N, 5, 1)
    W = N - 2 * e
    v = np.random.default_rng(8).standard_normal((W, W))
    a_x = np.zeros((N, N)); a_y = np.zeros((N, N))
    a_x[e:N - e, e:N - e] = v * np.gradient(f, axis=0)[e:N - e, e:N - e]
    a_y[e:N - e, e:N - e] = v * np.gradient(f, axis=1)[e:N - e, e:N - e]
    c = kw // 2

    def conv_form(Kx, Ky):
        return (scipy.signal.convolve2d(a_x, Kx, mode="full") +
                scipy.signal.convolve2d(a_y, Ky, mode="full"))[c:c + N, c:c + N]

    Kx, Ky = _kernel(orc, kw)
    np.testing.assert_allclose(orc.adjoint_g(v, f, e, Kx, Ky), -conv_form(Kx, Ky), rtol=1e-11,
                               atol=1e-13)
    yy, xx = np.meshgrid(np.arange(kw) - c, np.arange(kw) - c)
    Gs = np.exp(-(xx ** 2 + yy ** 2) / 8.0)
    Gs /= Gs.sum()
    np.testing.assert_allclose(orc.adjoint_g(v, f, e, Gs, Gs), conv_form(Gs, Gs), rtol=1e-11,
                               atol=1e-13)


# ---- tangent / tangent-adjoint ----------------------------------------------

def test_tangent_exact_central_difference(orc):
    N, kw, e = 38, 11, 6
    Kx, Ky = _kernel(orc, kw)
    f, g, df, dg = _rand_fields(N, 21, 4)
    t = orc.tangent(f, g, df, dg, e, Kx, Ky)
    for eps in (1e-1, 1.0):
        cd = (orc.forward(f + eps * df, g + eps * dg, e, Kx, Ky) -
              orc.forward(f - eps * df, g - eps * dg, e, Kx, Ky)) / (2 * eps)
        np.testing.assert_allclose(t, cd, rtol=1e-9, atol=1e-12)
    # dg-only and df-only modes
    z = np.zeros_like(f)
    np.testing.assert_allclose(orc.tangent(f, g, z, dg, e, Kx, Ky), orc.forward(f, dg, e, Kx, Ky),
                               rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(orc.tangent(f, g, df, z, e, Kx, Ky), orc.forward(df, g, e, Kx, Ky),
                               rtol=1e-12, atol=1e-14)


def test_tangent_adjoint_exact_central_difference(orc):
    N, kw, e = 36, 9, 5
    Kx, Ky = _kernel(orc, kw)
    f, g, df, dg = _rand_fields(N, 31, 4)
    W = N - 2 * e
    rng = np.random.default_rng(4)
    lam, dlam = rng.standard_normal((W, W)), rng.standard_normal((W, W))
    a, b = orc.tangent_adjoint(f, g, df, dg, lam, dlam, e, Kx, Ky)
    eps = 0.5
    pf, pg = orc.adjoint(lam + eps * dlam, f + eps * df, g + eps * dg, e, Kx, Ky)
    mf, mg = orc.adjoint(lam - eps * dlam, f - eps * df, g - eps * dg, e, Kx, Ky)
    np.testing.assert_allclose(a, (pf - mf) / (2 * eps), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(b, (pg - mg) / (2 * eps), rtol=1e-9, atol=1e-12)


# ---- CD metric ---------------------------------------------------------------

def test_cd_of_gaussian_pulse_is_fwhm():
    from oracle import cd
    N, sigma = 200, 6.0
    x = np.arange(N) - 0.5 * (N - 1)
    sig = np.exp(-x ** 2 / (2 * sigma ** 2))[None, :]
    left, right, w = cd.measure_cd(sig, threshold=0.5)
    fwhm = 2 * sigma * math.sqrt(2 * math.log(2))
    assert abs(w[0] - fwhm) < 2e-3
    assert abs((left[0] + right[0]) / 2 - 0.5 * (N - 1)) < 1e-6


def test_cd_upsample_matches_direct_sum():
    from oracle import cd
    rng = np.random.default_rng(2)
    s = rng.standard_normal((2, 23))
    L, A = 16, 8
    h = cd.polyphase_table(L, A)
    up = cd.upsample(s, L, A)
    for b in range(2):
        for m in (0, 5, 17, 100, 367):
            k, r = divmod(m, L)
            ref = sum(s[b, k + i - A] * h[r, i] for i in range(2 * A + 1) if 0 <= k + i - A < 23)
            assert abs(up[b, m] - ref) < 1e-12
    np.testing.assert_allclose(h.sum(axis=1), 1.0, atol=1e-12)
    np.testing.assert_allclose(up[:, ::L], s, atol=1e-9)   # phase 0 is the identity


def test_window_forms_match_whole_image(orc):
    N, kw, e = 44, 11, 6
    Kx, Ky = _kernel(orc, kw)
    f, g = _rand_fields(N, 77)
    W = N - 2 * e
    v = np.random.default_rng(6).standard_normal((W, W))
    h = orc.forward(f, g, e, Kx, Ky, s=1.3)
    np.testing.assert_allclose(orc.forward_window(f, g, e, Kx, Ky, 3, 5, 10, 7, s=1.3),
                               h[3:13, 5:12], rtol=1e-12, atol=1e-14)
    lf, lg = orc.adjoint(v, f, g, e, Kx, Ky, s=1.3)
    for (x0, y0, wx, wy) in ((0, 0, 9, 8), (4, 30, 12, 14), (35, 36, 9, 8)):
        a, b = orc.adjoint_window(v, f, g, e, Kx, Ky, x0, y0, wx, wy, s=1.3)
        np.testing.assert_allclose(a, lf[x0:x0 + wx, y0:y0 + wy], rtol=1e-11, atol=1e-14)
        np.testing.assert_allclose(b, lg[x0:x0 + wx, y0:y0 + wy], rtol=1e-11, atol=1e-14)


def test_tangent_window_forms_match_whole_image(orc):
    """The window forms of the tangent pair against the whole-image oracle
    tangent / tangent-adjoint (themselves pinned by exact central differences
    above); a swapped df / dg or lam_h / dlam_h changes every window."""
    N, kw, e = 40, 9, 5
    Kx, Ky = _kernel(orc, kw)
    f, g, df, dg = _rand_fields(N, 91, 4)
    W = N - 2 * e
    rng = np.random.default_rng(9)
    lam, dlam = rng.standard_normal((W, W)), rng.standard_normal((W, W))
    dh = orc.tangent(f, g, df, dg, e, Kx, Ky, s=0.7)
    np.testing.assert_allclose(orc.tangent_window(f, g, df, dg, e, Kx, Ky, 2, 4, 11, 9, s=0.7),
                               dh[2:13, 4:13], rtol=1e-11, atol=1e-14)
    a, b = orc.tangent_adjoint(f, g, df, dg, lam, dlam, e, Kx, Ky, s=0.7)
    for (x0, y0, wx, wy) in ((0, 0, 8, 9), (13, 25, 12, 15)):
        ta, tb = orc.tangent_adjoint_window(f, g, df, dg, lam, dlam, e, Kx, Ky, x0, y0, wx, wy,
                                            s=0.7)
        np.testing.assert_allclose(ta, a[x0:x0 + wx, y0:y0 + wy], rtol=1e-11, atol=1e-14)
        np.testing.assert_allclose(tb, b[x0:x0 + wx, y0:y0 + wy], rtol=1e-11, atol=1e-14)
    sa, _ = orc.tangent_adjoint_window(f, g, df, dg, dlam, lam, e, Kx, Ky, 13, 25, 12, 15, s=0.7)
    assert np.abs(sa - a[13:25, 25:40]).max() > 1e-3 * np.abs(a).max()


def test_cd_missing_edges_follow_reference_start_values():
    """A signal without a crossing on one side keeps the reference's search start
    value -/+FLT_MAX (cd_measure_util.cu:37,218-219), divided by L in FP32."""
    from oracle import cd
    N = 64
    x = np.arange(N)
    sig = np.stack([np.ones(N), np.where(x < 50, -1.0, 1.0), np.where(x < 20, 1.0, -1.0)])
    left, right, w = cd.measure_cd(sig, L=16, A=8)
    fm = float(np.finfo(np.float32).max)
    assert left[0] == -fm / 16 and right[0] == fm / 16 and w[0] == fm / 8
    assert left[1] == -fm / 16 and abs(right[1] - 49.5) < 0.6
    assert right[2] == fm / 16 and abs(left[2] - 19.5) < 0.6
    assert list(cd.found(left)) == [False, False, True]
    assert list(cd.found(w)) == [False, False, False]
