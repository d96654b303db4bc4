"""Ragged batched differentiation at full size: 64 instances, N in
{1024, 2048, 4096}, KW in {31, 61, 91, 121} (seed 4242), forward + adjoint +
tangent-adjoint for all groups captured in one graph.  Checks: graph replay
equals the eager run bit for bit; FP64-oracle windows for one instance of each
kernel width; the reduced-precision forms (FP16 MFMA stages, bfloat16 exchange)
against the FP32 form at the fp16 tolerance (rel 1e-2, abs 1e-3) on every
output."""
import numpy as np
import pytest

import shrinkage_full_tests_amd as srk
from shrinkage_full_tests_amd import ragged

pytestmark = pytest.mark.gpu


def test_ragged_64_instances_graph_oracle_and_mixed(orc):
    import torch
    rb = ragged.RaggedBatch()
    assert len(rb.inst) == 64 and len(rb.groups) <= 12
    g = rb.capture()
    eager = rb.outputs()
    for s in rb.groups:
        for k in eager[0]:
            s[k].zero_()
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    for s, ref in zip(rb.groups, eager):
        for k, r in ref.items():
            assert torch.equal(s[k], r), (s["N"], s["kw"], k)
    # FP64 oracle on windows, one group per kernel width
    from test_gpu_large import check_adjoint_window, check_forward_window
    seen = set()
    for s in rb.groups:
        if s["kw"] in seen:
            continue
        seen.add(s["kw"])
        _, _, hx, hy = srk.kernel_spectra_srsrk(s["kw"], s["e"] / 3.0, s["M"], with_spatial=True)
        Kx, Ky = hx.numpy().astype(np.float64), hy.numpy().astype(np.float64)
        e, W, N = s["e"], s["W"], s["N"]
        check_forward_window(orc, s["f"][0], s["g"][0], s["h"][0], e, Kx, Ky, W // 2, 5, 24)
        check_adjoint_window(orc, s["f"][0], s["g"][0], s["v"][0], s["lf"][0], s["lg"][0], e, Kx,
                             Ky, N - 30, N // 3, 24)
    # reduced-precision forms against the FP32 form
    for prec in ("mixed_fp16", "bf16_exchange"):
        rb.set_precision(prec)
        rb.run()
        torch.cuda.synchronize()
        for s, ref in zip(rb.groups, eager):
            for k, r in ref.items():
                a = s[k].double()
                n = r.double()
                lim = torch.clamp(1e-2 * torch.maximum(a.abs(), n.abs()), min=1e-3)
                bad = int(((a - n).abs() > lim).sum())
                assert bad == 0, (prec, s["N"], s["kw"], k, bad, float((a - n).abs().max()))
