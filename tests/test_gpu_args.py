"""Tensor arguments are checked on the host before any launch: a host tensor, a
wrong dtype, a non-contiguous view or an output too small for the call raises
ValueError instead of reaching a kernel as a wrong or host address."""
import pytest

import shrinkage_full_tests_amd as srk

pytestmark = pytest.mark.gpu

N, KW, M = 256, 31, 256
E_ = (KW - 1) // 2
W = N - 2 * E_


@pytest.fixture(scope="module")
def setup():
    import torch
    gx, gy = srk.kernel_spectra_srsrk(KW, 5.0, M)
    ws = srk.Workspace(N, W, E_, M, B_max=2)
    E = torch.rand((2, N, N), device="cuda")
    v = torch.randn((2, W, W), device="cuda")
    return torch, gx, gy, ws, E, v


def test_good_arguments_run(setup):
    torch, gx, gy, ws, E, v = setup
    R = srk.forward_batched(E, E, gx, gy, E_, 1.0, ws)
    lf, lg = srk.adjoint_batched(v, E, E, gx, gy, E_, 1.0, ws)
    torch.cuda.synchronize()
    assert torch.isfinite(R).all() and torch.isfinite(lf).all() and torch.isfinite(lg).all()


def test_host_tensor_rejected(setup):
    torch, gx, gy, ws, E, v = setup
    with pytest.raises(ValueError, match="CUDA"):
        srk.forward_batched(E.cpu(), E, gx, gy, E_, 1.0, ws)


def test_wrong_dtype_rejected(setup):
    torch, gx, gy, ws, E, v = setup
    with pytest.raises(ValueError, match="dtype"):
        srk.forward_batched(E.double(), E, gx, gy, E_, 1.0, ws)
    with pytest.raises(ValueError, match="dtype"):
        srk.forward_batched(E, E, torch.view_as_real(gx).contiguous(), gy, E_, 1.0, ws)


def test_non_contiguous_rejected(setup):
    torch, gx, gy, ws, E, v = setup
    with pytest.raises(ValueError, match="contiguous"):
        srk.adjoint_batched(v.transpose(1, 2), E, E, gx, gy, E_, 1.0, ws)


def test_small_output_rejected(setup):
    torch, gx, gy, ws, E, v = setup
    R = torch.empty((1, W, W), device="cuda")      # room for one image of two
    with pytest.raises(ValueError, match="elements"):
        srk.forward_batched(E, E, gx, gy, E_, 1.0, ws, R)
    small = torch.empty((M, M // 4), dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError, match="elements"):
        srk.srkTrainingForward(E, E, small, gy, N, W, E_, 1.0, 2, ws, torch.empty_like(v))
