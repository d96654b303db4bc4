"""The verification harness (shrinkage_full_tests_amd.testing) on host
operators with known derivatives: dump formats byte for byte, comparison
statistics, and that each tester passes a correct derivative and fails a
planted mistake (stTestCore semantics, see the module docstring)."""
import numpy as np
import pytest

from shrinkage_full_tests_amd import testing as T


def test_dump_buffer_text_format_and_round_trip(tmp_path):
    a = np.array([[1.0, 2.5], [3.0, -4.0], [1e-9, 123456789.0]], np.float32)  # W=3, H=2
    p = tmp_path / "x.txt"
    T.dump_buffer(str(p), a, name="x")
    lines = p.read_text().splitlines()
    assert lines[:3] == ["# name: x", "# dtype: float32", "# shape: 3 2"]
    # row y lists columns x = 0..W-1 (column-major source)
    assert lines[3] == "%.8g %.8g %.8g" % (1.0, 3.0, float(np.float32(1e-9)))
    assert lines[4] == "%.8g %.8g %.8g" % (2.5, -4.0, float(np.float32(123456789.0)))
    back, meta = T.load_buffer(str(p))
    assert meta["name"] == "x" and meta["dtype"] == "float32" and meta["shape"] == (3, 2)
    assert np.allclose(back, a.astype(np.float64), rtol=1e-7)
    # numpy.loadtxt (what the reference notebooks call) reads H rows of W values
    assert np.allclose(np.loadtxt(str(p)), a.T, rtol=1e-7)


def test_compare_statistics():
    ref = np.zeros((4, 3))
    test = ref.copy()
    test[2, 1] = 0.5
    test[0, 0] = -0.25
    r = T.compare(ref, test, tolerance=0.4)
    assert r.max_error == 0.5 and r.max_pos == (2, 1) and r.num_elements == 12
    assert abs(r.l2_error - np.sqrt(0.25 + 0.0625)) < 1e-15
    assert not r.passed
    assert T.compare(ref, test, tolerance=0.5).passed


def _conv_valid(x, k):
    W, H = x.shape
    kw = k.shape[0]
    out = np.zeros((W - kw + 1, H - kw + 1))
    for i in range(kw):
        for j in range(kw):
            out += k[i, j] * x[i:i + W - kw + 1, j:j + H - kw + 1]
    return out


def _conv_valid_T(v, k, shape):
    x = np.zeros(shape)
    kw = k.shape[0]
    for i in range(kw):
        for j in range(kw):
            x[i:i + v.shape[0], j:j + v.shape[1]] += k[i, j] * v
    return x


def test_adjoint_tester_passes_true_adjoint_and_fails_wrong_one(tmp_path):
    rng = np.random.default_rng(1)
    k = rng.standard_normal((3, 3))
    x = rng.standard_normal((12, 10))
    fwd = lambda a: _conv_valid(a, k)
    good = lambda v, a: _conv_valid_T(v, k, a.shape)
    bad = lambda v, a: _conv_valid_T(v, k.T, a.shape)     # planted mistake: transposed kernel
    cfg = T.AdjointConfig(num_samples=40, epsilon=1e-3, seed=7)
    r = T.adjoint_test(fwd, good, x, cfg)
    assert r.passed and r.num_samples == 40 and r.abs_error_max < 1e-7
    assert not T.adjoint_test(fwd, bad, x, cfg).passed
    p = tmp_path / "px.txt"
    T.dump_adjoint_pixels(str(p), r, label="conv")
    lines = p.read_text().splitlines()
    assert lines[0] == "# label: conv"
    assert lines[1] == "# nx ny analytical numerical rel_error abs_error"
    assert len(lines) == 42
    nx, ny = (int(t) for t in lines[2].split()[:2])
    assert (nx, ny) == (r.pixels[0].nx, r.pixels[0].ny)


def test_tangent_testers_epsilon_sweep():
    rng = np.random.default_rng(2)
    k = rng.standard_normal((3, 3))
    f, g = rng.standard_normal((2, 9, 9))
    df, dg = rng.standard_normal((2, 9, 9))
    # bilinear h(f, g) = f_crop * conv(g, k): central differences are exact
    fwd = lambda f_, g_: f_[1:-1, 1:-1] * _conv_valid(g_, k)
    tan = lambda s, d: d[0][1:-1, 1:-1] * _conv_valid(s[1], k) + s[0][1:-1, 1:-1] * _conv_valid(d[1], k)
    r = T.tangent_test(fwd, tan, [f, g], [df, dg])
    assert r.passed and len(r.per_epsilon) == 10
    assert r.best_abs_error_max == min(e.abs_error_max for e in r.per_epsilon)
    wrong = lambda s, d: d[0][1:-1, 1:-1] * _conv_valid(s[1], k)   # planted: dg term dropped
    rw = T.tangent_test(fwd, wrong, [f, g], [df, dg])
    assert not rw.passed and rw.best_num_violations > 0
    # one direction only: dg = None holds g fixed
    r1 = T.tangent_test(fwd, lambda s, d: d[0][1:-1, 1:-1] * _conv_valid(s[1], k), [f, g], [df, None])
    assert r1.passed
    # the same criterion on an "adjoint" output
    adj = lambda f_, g_: _conv_valid_T(f_[1:-1, 1:-1] * g_[1:-1, 1:-1], k, (11, 11))
    tadj = lambda s, d: _conv_valid_T(d[0][1:-1, 1:-1] * s[1][1:-1, 1:-1]
                                      + s[0][1:-1, 1:-1] * d[1][1:-1, 1:-1], k, (11, 11))
    assert T.tangent_adjoint_test(adj, tadj, [f, g], [df, dg]).passed
