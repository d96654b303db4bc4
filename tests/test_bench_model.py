"""The measurement model bench.py prices its roofline on (CPU only): the
algorithmic bytes of one fwd+adj step, each pass's compulsory bytes, and the
HBM traffic read back from the committed rocprofv3 --pmc summaries."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_algorithmic_bytes_of_the_headline_step():
    N, e, B = 512, 60, 5000
    W = N - 2 * e
    # E = I: read f (4N^2) twice, write h (4W^2), read v (4W^2), write two N x N outputs
    assert bench.alg_bytes(N, W, B, True) == B * (16 * N * N + 8 * W * W)
    assert abs(bench.alg_bytes(N, W, B, True) / 1e9 - 27.12) < 0.01
    assert bench.alg_bytes(N, W, 1, False) - bench.alg_bytes(N, W, 1, True) == 8 * N * N


def test_pass_bytes_cover_every_pass_and_exceed_the_algorithmic_bytes():
    N, e = 512, 60
    W = N - 2 * e
    pb = bench.pass_bytes_per_tile(N, N, e)
    assert set(pb) == {"k_p1_cols", "k_p1_adj_src", "k_p2_conv", "k_p2_corr", "k_p3_fwd",
                       "k_p3_adj"}
    # one X1 plane (M x (M/2+1) complex) read by the conv pass, two u planes written
    H = N // 2 + 1
    assert pb["k_p2_conv"] == 8 * N * H + 2 * 8 * H * W
    # one step runs P1 and P2 conv twice (forward and adjoint), the others once
    step = 2 * pb["k_p1_cols"] + 2 * pb["k_p2_conv"] + sum(
        pb[k] for k in ("k_p1_adj_src", "k_p2_corr", "k_p3_fwd", "k_p3_adj"))
    assert step > bench.alg_bytes(N, W, 1, True)
    assert bench.pass_bytes_per_tile(1024, 512, e) is None   # whole-image tiles only


def test_pmc_traffic_summaries_parse():
    fp32 = bench.pmc_traffic(512, 5000, 121, 512, "fp32")
    bf16 = bench.pmc_traffic(512, 5000, 121, 512, "bf16_exchange")
    if fp32 is None or bf16 is None:
        pytest.skip("no committed traffic summary")
    # measured traffic lies between the algorithmic bytes and the exchange-heavy bound
    alg = bench.alg_bytes(512, 392, 5000, True)
    assert alg < bf16 < fp32 < 10 * alg


def test_reference_training_workspace_size():
    # 36.74 GB at N = 512, B = 5000; 5.52 GB at N = 256, B = 3000 (the reference's notebook
    # reports 6.63 GB in use there, cuFFT work areas included)
    assert bench.ref_train_ws_bytes(512, 5000) == 36741120000
    assert abs(bench.ref_train_ws_bytes(256, 3000) / 1e9 - 5.52) < 0.01


def test_nominal_flop_convention():
    # 5 n log2 n per complex FFT; a fwd+adj tile is 9 M-line FFT sweeps' worth
    # (fwd 3, adj 6) of M^2 log2 M plus the products: 45 M^2 log2 M + 18 M^2
    M = 512
    assert bench.nominal_flops_per_tile(M) == 45 * M * M * 9 + 18 * M * M
    assert bench.nominal_flops_per_tile(256) == 45 * 256 * 256 * 8 + 18 * 256 * 256
