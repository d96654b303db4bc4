"""Gaussian-process fixtures (srkGpGenerate) against an independent
restatement in this file: MT19937-64 (Matsumoto-Nishimura 64-bit
parameters), libstdc++'s generate_canonical<double, 53> and its Marsaglia
polar normal_distribution, and numpy FFTs for the spectral synthesis of
GpGenerator (third_party/include/stTestCore/gp_generator.h:95-201).
Agreement to float rounding pins the native generator's noise stream, the
covariance filter, the padding and the column-major crop."""
import math

import numpy as np
import pytest

from shrinkage_full_tests_amd import patterns

M64 = (1 << 64) - 1


class MT19937_64:
    def __init__(self, seed):
        self.mt = [0] * 312
        self.mt[0] = seed & M64
        for i in range(1, 312):
            p = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (p ^ (p >> 62)) + i) & M64
        self.i = 312

    def _twist(self):
        mt = self.mt
        for k in range(312):
            y = (mt[k] & 0xFFFFFFFF80000000) | (mt[(k + 1) % 312] & 0x7FFFFFFF)
            v = mt[(k + 156) % 312] ^ (y >> 1)
            if y & 1:
                v ^= 0xB5026F5AA96619E9
            mt[k] = v
        self.i = 0

    def __call__(self):
        if self.i >= 312:
            self._twist()
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & M64


def canonical(g):
    r = float(g()) / 18446744073709551616.0
    return r if r < 1.0 else math.nextafter(1.0, 0.0)


class Normal:
    def __init__(self):
        self.saved = None

    def __call__(self, g):
        if self.saved is not None:
            v, self.saved = self.saved, None
            return v
        while True:
            x = 2.0 * canonical(g) - 1.0
            y = 2.0 * canonical(g) - 1.0
            r2 = x * x + y * y
            if not (r2 > 1.0 or r2 == 0.0):
                break
        m = math.sqrt(-2.0 * math.log(r2) / r2)
        self.saved = x * m
        return y * m


def gp_restated(W, H, ell, seed, count):
    Wp, Hp = 1 << (W - 1).bit_length(), 1 << (H - 1).bit_length()
    dy = np.where(np.arange(Hp) <= Hp // 2, np.arange(Hp), np.arange(Hp) - Hp)[:, None]
    dx = np.where(np.arange(Wp) <= Wp // 2, np.arange(Wp), np.arange(Wp) - Wp)[None, :]
    c = np.exp(-0.5 * ((dx / ell) ** 2 + (dy / ell) ** 2))
    amp = np.sqrt(np.maximum(0.0, np.fft.fft2(c).real))
    g, nd = MT19937_64(seed), Normal()
    out = []
    for _ in range(count):
        w = np.array([nd(g) for _ in range(Wp * Hp)]).reshape(Hp, Wp)   # row-major grid
        x = np.fft.ifft2(amp * np.fft.fft2(w)).real
        out.append(x[:H, :W].T.astype(np.float32))   # column-major [x, y]
    return np.stack(out)


def test_mt19937_64_known_answer():
    # the C++ standard fixes the 10000th output of a default-seeded mt19937_64
    g = MT19937_64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042


@pytest.mark.parametrize("W,H,ell,seed,count", [(64, 64, 8.0, 42, 2), (100, 60, 5.0, 43, 1)])
def test_gp_fields_match_restatement(W, H, ell, seed, count):
    got = patterns.gp_fields(W, H, ell, seed, count=count)
    want = gp_restated(W, H, ell, seed, count)
    assert got.shape == want.shape == (count, W, H)
    assert np.abs(got - want).max() <= 2e-6 * np.abs(want).max()


def test_gp_field_statistics_and_seeds():
    a = patterns.gp_field(128, 8.0, 42)
    b = patterns.gp_field(128, 8.0, 43)
    assert not np.array_equal(a, b)
    assert np.array_equal(a, patterns.gp_field(128, 8.0, 42))
    # unit-variance stationary field with Gaussian covariance of length 8 px
    many = patterns.gp_fields(128, 128, 8.0, 7, count=8)
    v = many.var()
    lag = np.mean(many[:, 8:, :] * many[:, :-8, :]) / v
    assert 0.7 < v < 1.3 and abs(lag - math.exp(-0.5)) < 0.1
