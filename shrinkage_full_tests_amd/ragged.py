"""Ragged batched differentiation: a list of images of different sizes and
kernel widths run as uniform groups (one workspace and one spectrum pair per
(N, KW)), forward + adjoint + tangent-adjoint, eagerly or replayed from one
captured hipGraph.

The default instance list is the ragged configuration of the benchmark
suite: 64 instances with N drawn uniformly from {1024, 2048, 4096} and KW
from {31, 61, 91, 121} (seed 4242), Gaussian random fields, sigma = e / 3.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

import shrinkage_full_tests_amd as srk
from shrinkage_full_tests_amd import patterns


def instances(count: int = 64, seed: int = 4242, sizes: Sequence[int] = (1024, 2048, 4096),
              widths: Sequence[int] = (31, 61, 91, 121)) -> List[Tuple[int, int]]:
    rng = np.random.default_rng(seed)
    return [(int(rng.choice(sizes)), int(rng.choice(widths))) for _ in range(count)]


def tile_for(N: int, kw: int, precision: str = "fp32") -> int:
    """Tile size for a group: one whole-image tile up to N = 1024, else
    1024-point overlap-and-save tiles (2048 if the halo leaves no room).  The
    FP16-MFMA form runs faster on 512-point tiles beyond N = 512 (its
    1024-point transforms hold too many registers)."""
    e = (kw - 1) // 2
    if precision == "mixed_fp16" and N > 512 and 512 > 2 * e + 2:
        return 512
    if N <= 1024:
        return next(M for M in srk.SUPPORTED_M if M >= N)
    return 1024 if 1024 > 2 * e + 2 else 2048


class RaggedBatch:
    """Inputs, outputs, workspaces and spectra of a ragged instance list."""

    def __init__(self, inst: Optional[List[Tuple[int, int]]] = None, precision: str = "fp32",
                 seed: int = 7, tangent_adjoint: bool = True):
        import torch
        self.inst = inst if inst is not None else instances()
        self.tangent_adjoint = tangent_adjoint
        by_key: Dict[Tuple[int, int], List[int]] = {}
        for i, key in enumerate(self.inst):
            by_key.setdefault(key, []).append(i)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        self.groups = []
        for gi, ((N, kw), idx) in enumerate(sorted(by_key.items())):
            e = (kw - 1) // 2
            W = N - 2 * e
            B = len(idx)
            M = tile_for(N, kw, precision)
            gx, gy = srk.kernel_spectra_srsrk(kw, e / 3.0, M)
            ws = srk.Workspace(N, W, e, M, B_max=B, precision=precision, scratch_mb=1024)
            s = dict(N=N, kw=kw, e=e, W=W, B=B, M=M, idx=idx, gx=gx, gy=gy, ws=ws,
                     f=patterns.grf_torch(B, N, seed=1000 + 10 * gi),
                     g=patterns.grf_torch(B, N, seed=2000 + 10 * gi),
                     v=torch.randn((B, W, W), generator=gen, device="cuda"),
                     h=torch.empty((B, W, W), device="cuda"),
                     lf=torch.empty((B, N, N), device="cuda"),
                     lg=torch.empty((B, N, N), device="cuda"))
            if tangent_adjoint:
                s.update(df=patterns.grf_torch(B, N, seed=3000 + 10 * gi),
                         dg=patterns.grf_torch(B, N, seed=4000 + 10 * gi),
                         dv=torch.randn((B, W, W), generator=gen, device="cuda"),
                         tf=torch.empty((B, N, N), device="cuda"),
                         tg=torch.empty((B, N, N), device="cuda"))
            self.groups.append(s)
        self.graph = None

    @property
    def pixels(self) -> int:
        return sum(N * N for N, _ in self.inst)

    def set_precision(self, precision: str) -> None:
        """Switch every group to a precision form, rebuilding the workspace and
        spectra of groups whose tile size differs for that form."""
        for s in self.groups:
            M = tile_for(s["N"], s["kw"], precision)
            if M != s["M"]:
                s["ws"].close()
                s["gx"], s["gy"] = srk.kernel_spectra_srsrk(s["kw"], s["e"] / 3.0, M)
                s["ws"] = srk.Workspace(s["N"], s["W"], s["e"], M, B_max=s["B"],
                                        precision=precision, scratch_mb=1024)
                s["M"] = M
            s["ws"].set_precision(precision)
        self.graph = None

    def run(self) -> None:
        for s in self.groups:
            srk.forward_batched(s["g"], s["f"], s["gx"], s["gy"], s["e"], 1.0, s["ws"], s["h"])
            srk.adjoint_batched(s["v"], s["g"], s["f"], s["gx"], s["gy"], s["e"], 1.0, s["ws"],
                                s["lf"], s["lg"])
            if self.tangent_adjoint:
                srk.tangent_adjoint_batched(s["f"], s["g"], s["df"], s["dg"], s["v"], s["dv"],
                                            s["gx"], s["gy"], s["e"], 1.0, s["ws"], s["tf"],
                                            s["tg"])

    def capture(self):
        """Capture run() into one graph (after an eager warm-up run)."""
        import torch
        self.run()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        st = torch.cuda.Stream()
        st.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(st):
            with torch.cuda.graph(g, stream=st):
                self.run()
        torch.cuda.current_stream().wait_stream(st)
        self.graph = g
        return g

    def outputs(self) -> List[Dict]:
        keys = ("h", "lf", "lg") + (("tf", "tg") if self.tangent_adjoint else ())
        return [{k: s[k].clone() for k in keys} for s in self.groups]
