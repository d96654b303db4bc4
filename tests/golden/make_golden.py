"""Writes tests/golden/k1_scipy.json: FP64 golden values of the modified Bessel
function K1 from scipy.special.kv(1, x) (a library routine, independent of the
oracle's integral and of the product's series/integral), plus SRSRK Green's
tensor samples built from those values with the kernel formula of
training/main_cd_training.cu:63-77.

    python tests/golden/make_golden.py
"""
import json
import math
import os

import numpy as np
import scipy
import scipy.special as sp

HERE = os.path.dirname(os.path.abspath(__file__))


def srsrk(kw, sigma, dx=1.0, cw=3.0):
    gamma, Rw = 1.0 / sigma, cw * sigma
    coeff, c = gamma / (2 * math.pi), kw // 2
    Kx = np.zeros((kw, kw)); Ky = np.zeros((kw, kw))
    for i in range(kw):
        for j in range(kw):
            x, y = (i - c) * dx, (j - c) * dx
            r = math.hypot(x, y)
            if r < 1e-6 * dx or r > Rw:
                continue
            t = r / Rw
            w = 0.42 + 0.5 * math.cos(math.pi * t) + 0.08 * math.cos(2 * math.pi * t)
            base = coeff * sp.kv(1, gamma * r) * w
            Kx[i, j] = base * x / r
            Ky[i, j] = base * y / r
    return Kx, Ky


def main():
    xs = sorted(set([0.01, 0.05, 0.0625, 0.1, 0.2, 0.25, 0.5, 0.75, 1.0, 1.5, 1.999, 2.0, 2.001,
                     2.5, 3.0, 4.0, 5.0, 7.5, 10.0, 20.0] +
                    [round(0.05 * k, 6) for k in range(1, 61)]))
    k1 = [[x, float(sp.kv(1, x))] for x in xs]
    samples = {}
    for kw, sigma in ((31, 5.0), (61, 10.0), (121, 20.0)):
        Kx, Ky = srsrk(kw, sigma)
        c = kw // 2
        pts = [(c + 1, c), (c, c + 1), (c + 3, c - 2), (c - 7, c + 5), (c + c, c), (0, 0),
               (c + c // 2, c + c // 3)]
        samples[f"{kw}_{sigma}"] = [[i, j, Kx[i, j], Ky[i, j]] for i, j in pts]
    out = {"scipy": scipy.__version__, "k1": k1, "srsrk_samples": samples}
    with open(os.path.join(HERE, "k1_scipy.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
