"""One workspace used from two streams by one host thread: the second call
waits for the first (per-workspace completion event), so a call that imports
other spectra into the shared workspace cannot overwrite the first call's
spectra or exchange buffers while it still runs."""
import pytest

import shrinkage_full_tests_amd as srk
from shrinkage_full_tests_amd import patterns

pytestmark = pytest.mark.gpu


def test_calls_on_two_streams_are_ordered():
    import torch
    N, M, B, e = 512, 512, 64, 60
    W = N - 2 * e
    ga = srk.kernel_spectra_srsrk(121, 20.0, M)
    gb = srk.kernel_spectra_srsrk(61, 10.0, M)
    E = patterns.grf_torch(B, N, seed=31)
    I = patterns.grf_torch(B, N, seed=32)
    ws = srk.Workspace(N, W, e, M, B_max=B)
    ra = srk.forward_batched(E, I, *ga, e, 1.0, ws)
    rb = srk.forward_batched(I, E, *gb, e, 1.0, ws)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):
        oa = torch.empty_like(ra)
        ob = torch.empty_like(rb)
        with torch.cuda.stream(s1):
            srk.forward_batched(E, I, *ga, e, 1.0, ws, oa, stream=s1)
        with torch.cuda.stream(s2):
            srk.forward_batched(I, E, *gb, e, 1.0, ws, ob, stream=s2)
        torch.cuda.synchronize()
        assert torch.equal(oa, ra) and torch.equal(ob, rb)
