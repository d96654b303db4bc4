"""Device-generated SRSRK spectra (srkKernelSpectraSRSRKDevice) against the
host FP64 path of srkKernelSpectraSRSRK, which tests/test_abi.py pins to scipy
(kv(1, x)) and to the notebook known answer.  Both evaluate every step in
FP64 and round once to FP32, so they agree to within one FP32 rounding."""
import numpy as np
import pytest

import shrinkage_full_tests_amd as srk

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kw,sigma,M", [(31, 5.0, 256), (61, 10.0, 512), (121, 20.0, 512),
                                        (91, 15.0, 1024), (121, 20.0, 2048)])
def test_device_spectra_match_host(kw, sigma, M):
    import torch
    gx_h, gy_h, hx, hy = srk.kernel_spectra_srsrk(kw, sigma, M, with_spatial=True)
    gx_d, gy_d, kx_d, ky_d = srk.kernel_spectra_srsrk_device(kw, sigma, M, with_spatial=True)
    torch.cuda.synchronize()
    for d, h in ((gx_d, gx_h), (gy_d, gy_h)):
        d = torch.view_as_real(d).cpu().numpy().astype(np.float64)
        h = torch.view_as_real(h).cpu().numpy().astype(np.float64)
        scale = np.abs(h).max()
        assert scale > 0
        assert np.abs(d - h).max() <= 1e-6 * scale, np.abs(d - h).max() / scale
    for d, h in ((kx_d, hx), (ky_d, hy)):
        d = d.cpu().numpy().astype(np.float64)
        h = h.numpy().astype(np.float64)
        assert np.abs(d - h).max() <= 1e-6 * np.abs(h).max()


def test_forward_with_device_spectra(orc):
    import torch
    N, kw, e = 256, 31, 15
    W = N - 2 * e
    gx, gy = srk.kernel_spectra_srsrk_device(kw, 5.0, N)
    _, _, hx, hy = srk.kernel_spectra_srsrk(kw, 5.0, N, with_spatial=True)
    from shrinkage_full_tests_amd import patterns
    f = patterns.grf(N, 77)[0]
    ws = srk.Workspace(N, W, e, N)
    fd = torch.from_numpy(np.ascontiguousarray(f[None], np.float32)).cuda()
    h = srk.forward_batched(fd, fd, gx, gy, e, 1.0, ws)
    ref = orc.forward(f, f, e, hx.numpy().astype(np.float64), hy.numpy().astype(np.float64))
    got = h.cpu().numpy()[0].astype(np.float64)
    lim = np.maximum(1e-5, 1e-3 * np.maximum(np.abs(got), np.abs(ref)))
    assert (np.abs(got - ref) <= lim).all()


def test_device_spectra_argument_errors():
    with pytest.raises(srk.SrkError):
        srk.kernel_spectra_srsrk_device(30, 5.0, 256)     # even width
    with pytest.raises(srk.SrkError):
        srk.kernel_spectra_srsrk_device(31, -1.0, 256)    # sigma <= 0
    with pytest.raises(srk.SrkError):
        srk.kernel_spectra_srsrk_device(31, 5.0, 16)      # size below the kernel width


def test_device_spectra_any_size_match_host():
    """Spectra at a size that is not a power of two (the reference plans cuFFT
    at any N) agree between the device and the host FP64 paths."""
    import torch
    gx_h, gy_h = srk.kernel_spectra_srsrk(31, 5.0, 300)
    gx_d, gy_d = srk.kernel_spectra_srsrk_device(31, 5.0, 300)
    torch.cuda.synchronize()
    for h, d in ((gx_h, gx_d), (gy_h, gy_d)):
        peak = float(h.abs().max())
        assert float((h - d).abs().max()) <= 1e-6 * peak
