"""The stTestCore-style testers (shrinkage_full_tests_amd.testing) applied to
the HIP operators through the C ABI, as the reference drives its kernels:
AdjointTester on lambda_f and lambda_g (q = sum h^2, adjoint_tester.h:28-140),
TangentTester on the tangent (tangent_tester.h:27-62) and
TangentAdjointTester on the tangent-adjoint (tangent_adjoint_tester.h:28-38),
at the fp32 tolerance.  h is bilinear in (f, g), so central differences carry
rounding error only."""
import numpy as np
import pytest

import shrinkage_full_tests_amd as srk
from shrinkage_full_tests_amd import patterns
from shrinkage_full_tests_amd import testing as T

pytestmark = pytest.mark.gpu

N, KW, M = 256, 31, 256
E = (KW - 1) // 2
W = N - 2 * E


@pytest.fixture(scope="module")
def setup():
    import torch
    gx, gy = srk.kernel_spectra_srsrk(KW, 5.0, M)
    ws = srk.Workspace(N, W, E, M)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    f = dev(patterns.grf(N, 501)[0])
    g = dev(patterns.grf(N, 502)[0])
    rng = np.random.default_rng(503)
    df = dev(rng.standard_normal((N, N)))
    dg = dev(rng.standard_normal((N, N)))
    lam = dev(rng.standard_normal((W, W)))
    dlam = dev(rng.standard_normal((W, W)))
    return gx, gy, ws, f, g, df, dg, lam, dlam


def _fwd(gx, gy, ws):
    return lambda f_, g_: srk.forward_batched(g_[None], f_[None], gx, gy, E, 1.0, ws)[0]


def test_adjoint_tester_lambda_f_and_lambda_g(setup):
    gx, gy, ws, f, g, *_ = setup
    fwd = _fwd(gx, gy, ws)
    adj = lambda v, f_, g_: srk.adjoint_batched(v[None], g_[None], f_[None], gx, gy, E, 1.0, ws)
    cfg = T.AdjointConfig(num_samples=48, epsilon=1e-1, seed=11)
    rf = T.adjoint_test(lambda x: fwd(x, g), lambda v, x: adj(v, x, g)[0][0], f, cfg)
    assert rf.passed, (rf.abs_error_max, rf.rel_error_max)
    rg = T.adjoint_test(lambda x: fwd(f, x), lambda v, x: adj(v, f, x)[1][0], g, cfg)
    assert rg.passed, (rg.abs_error_max, rg.rel_error_max)


def test_tangent_tester(setup):
    gx, gy, ws, f, g, df, dg, *_ = setup
    fwd = _fwd(gx, gy, ws)

    def tan(s, d):
        return srk.tangent_batched(s[0][None], s[1][None], d[0][None], d[1][None], gx, gy, E,
                                   1.0, ws)[0]
    r = T.tangent_test(fwd, tan, [f, g], [df, dg],
                       T.TangentConfig(num_samples=4096, seed=12, epsilons=[1e-2, 1e-1]))
    assert r.passed, (r.best_abs_error_max, r.best_num_violations)


def test_tangent_adjoint_tester(setup):
    gx, gy, ws, f, g, df, dg, lam, dlam = setup

    def adj_out(which):
        return lambda f_, g_, l_: srk.adjoint_batched(l_[None], g_[None], f_[None], gx, gy, E,
                                                      1.0, ws)[which][0]

    def tadj(which):
        return lambda s, d: srk.tangent_adjoint_batched(
            s[0][None], s[1][None], d[0][None], d[1][None], s[2][None], d[2][None], gx, gy, E,
            1.0, ws)[which][0]
    cfg = T.TangentConfig(num_samples=4096, seed=13, epsilons=[1e-2, 1e-1])
    for which in (0, 1):
        r = T.tangent_adjoint_test(adj_out(which), tadj(which), [f, g, lam], [df, dg, dlam], cfg)
        assert r.passed, (which, r.best_abs_error_max, r.best_num_violations)
