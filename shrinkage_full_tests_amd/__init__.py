"""MI355X-native SRK shrinkage operator (gfx950 HIP kernels behind a C ABI).

Host-side mirror of the reference operator interface.  The entry points keep
the reference names and argument meaning (see include/srk.h); arrays are
torch tensors on the GPU in the reference's column-major layout, i.e. a
contiguous tensor of shape [B, N, N] indexed [b, column, row].

Nothing here computes on the CPU: every operator runs the HIP library
libsrk_amd.so and raises if it is not built.
"""
from __future__ import annotations

import ctypes

from ._lib import SrkError, call, lib  # noqa: F401

__all__ = [
    "SrkError", "lib", "Workspace", "kernel_spectra_srsrk", "kernel_spectra_srsrk_device",
    "precompute_kernel_freq",
    "srkTrainingForward", "srkForwardOaS", "adjointSRK", "shrinkageTangentOaS",
    "shrinkageTangentAdjointOaS", "forward_batched", "adjoint_batched", "tangent_batched",
    "forward_batched_u", "adjoint_batched_u",
    "tangent_adjoint_batched", "ShrinkageFunction", "erosion_for", "SUPPORTED_M",
    "measure_cd_batched", "shrinkage_cd_batched", "check_spectra_support",
    "build_polyphase_table", "upsample_batched", "measure_cd_batched_ex", "cd_edge_found",
]

SUPPORTED_M = (256, 512, 1024, 2048)
KERNEL_KINDS = 7  # SRK_KERNEL_KINDS
# SRK_PREC_FP32, SRK_PREC_MIXED_FP16, SRK_PREC_BF16_EXCHANGE
PRECISIONS = {"fp32": 0, "mixed_fp16": 1, "bf16_exchange": 2}


def _torch():
    import torch
    return torch


def _ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


def _dev(t, numel: int, what: str, dtype: str = "float32", optional: bool = False) -> int:
    """Device address of a tensor handed to the library: a contiguous CUDA
    tensor of `dtype` with at least `numel` elements.  Anything else raises
    ValueError here instead of reaching a kernel as a wrong or host address."""
    torch = _torch()
    if t is None:
        if optional:
            return 0
        raise ValueError(f"{what}: a tensor is required")
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{what}: needs a CUDA tensor")
    if t.dtype != getattr(torch, dtype):
        raise ValueError(f"{what}: dtype {t.dtype}, expected torch.{dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{what}: needs a contiguous tensor")
    if t.numel() < numel:
        raise ValueError(f"{what}: {t.numel()} elements, needs {numel}")
    return int(t.data_ptr())


def _spec(G, ws, what: str) -> int:
    """Reference-layout spectrum [M, M/2+1] complex64 at the workspace's tile size."""
    return _dev(G, ws.M * (ws.M // 2 + 1), what, "complex64")


def _stream(stream) -> int:
    torch = _torch()
    if stream is None:
        return int(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


def erosion_for(kw: int) -> int:
    """Reference default erosion (KW-1)/2 (README.md 'Derived parameters')."""
    return (kw - 1) // 2


class Workspace:
    """Opaque device workspace for images of size N, erosion e, batches <= B_max,
    FFT/tile size M (the size of the spectra handed to the calls)."""

    def __init__(self, N: int, W: int, erosion: int, M: int, B_max: int = 1, scratch_mb: int = 0,
                 precision="fp32"):
        h = ctypes.c_void_p()
        call("srkWorkspaceCreate", N, W, erosion, M, B_max, scratch_mb, ctypes.byref(h))
        self.handle = h.value
        self.N, self.W, self.erosion, self.M, self.B_max = N, W, erosion, M, B_max
        self.precision = "fp32"
        if precision != "fp32":
            self.set_precision(precision)

    def close(self):
        if self.handle:
            call("srkWorkspaceDestroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def nbytes(self) -> int:
        return int(lib().srkWorkspaceBytes(self.handle))

    def set_chunk(self, tiles: int) -> None:
        call("srkWorkspaceSetChunk", self.handle, int(tiles))

    def set_tuning(self, chunk_tiles=None, p2_rows=None, concurrency=None, fused=None,
                   fused_slots=None, fused_lag=None) -> None:
        """Autotuner knobs: tiles per launch chunk, k_y bands per row-pass block,
        chunks alternating between two streams (1, the default) or one (0), the
        fused one-launch pipeline for one-tile images (1, or 0, the default), and
        its exchange-ring slots per XCD and stage lag in rounds (lag < slots <= 4)."""
        if chunk_tiles is not None:
            call("srkWorkspaceSetTuning", self.handle, 0, int(chunk_tiles))
        if p2_rows is not None:
            call("srkWorkspaceSetTuning", self.handle, 1, int(p2_rows))
        if concurrency is not None:
            call("srkWorkspaceSetTuning", self.handle, 2, int(concurrency))
        if fused is not None:
            call("srkWorkspaceSetTuning", self.handle, 3, int(bool(fused)))
        if fused_slots is not None and fused_lag is not None:
            call("srkWorkspaceSetTuning", self.handle, 5, 1)  # below any slot count
        if fused_slots is not None:
            call("srkWorkspaceSetTuning", self.handle, 4, int(fused_slots))
        if fused_lag is not None:
            call("srkWorkspaceSetTuning", self.handle, 5, int(fused_lag))

    def fused_status(self):
        """(available, timed-out waits of the last fused launch); synchronises."""
        av, to = ctypes.c_int(), ctypes.c_int()
        st = lib().srkWorkspaceFusedStatus(self.handle, ctypes.byref(av), ctypes.byref(to))
        if st not in (0, -8):
            raise SrkError(st, "srkWorkspaceFusedStatus")
        return bool(av.value), int(to.value)

    def set_precision(self, precision) -> None:
        """"fp32" (default), "mixed_fp16" (FP16 MFMA DFT stages, FP32 accumulation) or
        "bf16_exchange" (FP32 transforms, bfloat16 spectra between the passes except
        the adjoints' recomputed convolution)."""
        code = PRECISIONS.get(precision, precision)
        call("srkWorkspaceSetPrecision", self.handle, int(code))
        self.precision = precision

    def set_timing(self, enable: bool = True) -> None:
        """Bracket every later engine launch on this workspace with HIP events."""
        call("srkWorkspaceSetTiming", self.handle, int(bool(enable)))

    def kernel_times(self) -> dict:
        """Wait for the timed launches; {kernel: (total ms, launches, tiles)}, then reset."""
        n = KERNEL_KINDS
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int * n)()
        tiles = (ctypes.c_longlong * n)()
        call("srkWorkspaceKernelTimes", self.handle, ctypes.addressof(ms), ctypes.addressof(cnt),
             ctypes.addressof(tiles))
        return {lib().srkKernelKindName(k).decode(): (ms[k], cnt[k], tiles[k]) for k in range(n)}

    def info(self) -> dict:
        v = [ctypes.c_int() for _ in range(5)]
        call("srkWorkspaceInfo", self.handle, *[ctypes.byref(x) for x in v])
        keys = ("M", "fwd_tiles_per_image", "adj_tiles_per_image", "chunk_fwd", "chunk_adj")
        return {k: x.value for k, x in zip(keys, v)}


def kernel_spectra_srsrk(kw: int, sigma: float, M: int, dx: float = 1.0, cw: float = 3.0,
                         device="cuda", with_spatial: bool = False):
    """FP64-built SRSRK spectra (reference layout, complex64 [M, M/2+1]) at tile
    size M; cached inside the library by (kw, sigma, dx, cw, M)."""
    torch = _torch()
    gx = torch.empty((M, M // 2 + 1), dtype=torch.complex64, device=device)
    gy = torch.empty_like(gx)
    hx = hy = None
    if with_spatial:
        hx = torch.empty((kw, kw), dtype=torch.float32)
        hy = torch.empty((kw, kw), dtype=torch.float32)
    call("srkKernelSpectraSRSRK", kw, float(sigma), float(dx), float(cw), M, _ptr(gx), _ptr(gy),
         _ptr(hx), _ptr(hy), _stream(None))
    if with_spatial:
        return gx, gy, hx, hy
    return gx, gy


def kernel_spectra_srsrk_device(kw: int, sigma: float, M: int, dx: float = 1.0, cw: float = 3.0,
                                with_spatial: bool = False, stream=None):
    """SRSRK spectra computed in FP64 on the device (no host round trip);
    with_spatial also returns the [kw, kw] spatial kernels as device tensors."""
    torch = _torch()
    gx = torch.empty((M, M // 2 + 1), dtype=torch.complex64, device="cuda")
    gy = torch.empty_like(gx)
    hx = hy = None
    if with_spatial:
        hx = torch.empty((kw, kw), dtype=torch.float32, device="cuda")
        hy = torch.empty_like(hx)
    call("srkKernelSpectraSRSRKDevice", kw, float(sigma), float(dx), float(cw), M, _ptr(gx),
         _ptr(gy), _ptr(hx), _ptr(hy), _stream(stream))
    return (gx, gy, hx, hy) if with_spatial else (gx, gy)


def precompute_kernel_freq(G, M: int, device="cuda"):
    """srkPrecomputeKernelFreq of a host [KW, KW] float32 kernel (indexed [x, y])."""
    torch = _torch()
    G = G.detach().to("cpu", torch.float32).contiguous()
    out = torch.empty((M, M // 2 + 1), dtype=torch.complex64, device=device)
    call("srkPrecomputeKernelFreq", _ptr(G), G.shape[0], G.shape[1], M, _ptr(out), _stream(None))
    return out


# ---- drop-in entry points (reference names) --------------------------------

def srkTrainingForward(E, I, Gx_freq, Gy_freq, N, W, erosion, dx, B, ws, R, stream=None):
    """Batched forward with spectra at size N (reference layout [N, N/2+1]),
    for any N; the workspace's tile size M may differ from N."""
    img, val = B * N * N, B * W * W
    spec = N * (N // 2 + 1)
    call("srkTrainingForward", _dev(E, img, "E"), _dev(I, img, "I"),
         _dev(Gx_freq, spec, "Gx", "complex64"), _dev(Gy_freq, spec, "Gy", "complex64"), N, W,
         erosion, float(dx), B, _stream(stream), _dev(R, val, "R"), ws.handle)
    return R


def check_spectra_support(Gx_freq, Gy_freq, Ns, erosion, ws, stream=None) -> float:
    """srkCheckSpectraSupport: relative RMS of the size-Ns spectra's spatial
    kernel outside the erosion box; raises SrkError (SRK_ERR_KERNEL) above 1e-4."""
    r = ctypes.c_double(-1.0)
    spec = Ns * (Ns // 2 + 1)
    call("srkCheckSpectraSupport", _dev(Gx_freq, spec, "Gx", "complex64"),
         _dev(Gy_freq, spec, "Gy", "complex64"), Ns, erosion, _stream(stream), ws.handle,
         ctypes.byref(r))
    return r.value


def srkForwardOaS(E, I, Gx_freq, Gy_freq, N, W, erosion, dx, ws, R, ux=None, uy=None,
                  stream=None):
    img, val = N * N, W * W
    call("srkForwardOaS", _dev(E, img, "E"), _dev(I, img, "I"), _spec(Gx_freq, ws, "Gx"),
         _spec(Gy_freq, ws, "Gy"), N, W, erosion, float(dx), _stream(stream), _dev(R, val, "R"),
         ws.handle, _dev(ux, val, "ux", optional=True), _dev(uy, val, "uy", optional=True))
    return R


def adjointSRK(v, ux, uy, I, Gx_freq, Gy_freq, N, W, erosion, dx, ws, dLdI, dLdE, stream=None):
    img, val = N * N, W * W
    call("adjointSRK", _dev(v, val, "v"), _dev(ux, val, "ux"), _dev(uy, val, "uy"),
         _dev(I, img, "I"), _spec(Gx_freq, ws, "Gx"), _spec(Gy_freq, ws, "Gy"), N, W, erosion,
         float(dx), _stream(stream), _dev(dLdI, img, "dLdI"), _dev(dLdE, img, "dLdE"),
         ws.handle)
    return dLdI, dLdE


def shrinkageTangentOaS(f, g, df, dg, Gx_freq, Gy_freq, N, W, erosion, dx, ws, dh, stream=None):
    img, val = N * N, W * W
    call("shrinkageTangentOaS", _dev(f, img, "f"), _dev(g, img, "g"), _dev(df, img, "df"),
         _dev(dg, img, "dg"), _spec(Gx_freq, ws, "Gx"), _spec(Gy_freq, ws, "Gy"), N, W, erosion,
         float(dx), _stream(stream), _dev(dh, val, "dh"), ws.handle)
    return dh


def shrinkageTangentAdjointOaS(f, g, df, dg, lam_h, dlam_h, Gx_freq, Gy_freq, N, W, erosion, dx,
                               ws, dLam_f, dLam_g, stream=None):
    img, val = N * N, W * W
    call("shrinkageTangentAdjointOaS", _dev(f, img, "f"), _dev(g, img, "g"), _dev(df, img, "df"),
         _dev(dg, img, "dg"), _dev(lam_h, val, "lam_h"), _dev(dlam_h, val, "dlam_h"),
         _spec(Gx_freq, ws, "Gx"), _spec(Gy_freq, ws, "Gy"), N, W, erosion, float(dx),
         _stream(stream), _dev(dLam_f, img, "dLam_f"), _dev(dLam_g, img, "dLam_g"), ws.handle)
    return dLam_f, dLam_g


# ---- batched entry points ---------------------------------------------------

def forward_batched(E, I, Gx_freq, Gy_freq, erosion, dx, ws, R=None, stream=None):
    torch = _torch()
    B, N = E.shape[0], E.shape[-1]
    W = N - 2 * erosion
    if R is None:
        R = torch.empty((B, W, W), dtype=torch.float32, device=E.device)
    img, val = B * N * N, B * W * W
    call("srkForwardBatched", _dev(E, img, "E"), _dev(I, img, "I"), _spec(Gx_freq, ws, "Gx"),
         _spec(Gy_freq, ws, "Gy"), N, W, erosion, float(dx), B, _stream(stream),
         _dev(R, val, "R"), ws.handle)
    return R


def adjoint_batched(v, E, I, Gx_freq, Gy_freq, erosion, dx, ws, dLdI=None, dLdE=None,
                    stream=None):
    torch = _torch()
    B, N = E.shape[0], E.shape[-1]
    W = N - 2 * erosion
    if dLdI is None:
        dLdI = torch.empty((B, N, N), dtype=torch.float32, device=E.device)
    if dLdE is None:
        dLdE = torch.empty((B, N, N), dtype=torch.float32, device=E.device)
    img, val = B * N * N, B * W * W
    call("srkAdjointBatched", _dev(v, val, "v"), _dev(E, img, "E"), _dev(I, img, "I"),
         _spec(Gx_freq, ws, "Gx"), _spec(Gy_freq, ws, "Gy"), N, W, erosion, float(dx), B,
         _stream(stream), _dev(dLdI, img, "dLdI"), _dev(dLdE, img, "dLdE"), ws.handle)
    return dLdI, dLdE


def forward_batched_u(E, I, Gx_freq, Gy_freq, erosion, dx, ws, R=None, ux=None, uy=None,
                      stream=None):
    """Batched forward that also stores u_x, u_y [B, W, W] for adjoint_batched_u."""
    torch = _torch()
    B, N = E.shape[0], E.shape[-1]
    W = N - 2 * erosion
    R, ux, uy = (t if t is not None else torch.empty((B, W, W), dtype=torch.float32,
                                                       device=E.device) for t in (R, ux, uy))
    img, val = B * N * N, B * W * W
    call("srkForwardBatchedU", _dev(E, img, "E"), _dev(I, img, "I"), _spec(Gx_freq, ws, "Gx"),
         _spec(Gy_freq, ws, "Gy"), N, W, erosion, float(dx), B, _stream(stream),
         _dev(R, val, "R"), _dev(ux, val, "ux"), _dev(uy, val, "uy"), ws.handle)
    return R, ux, uy


def adjoint_batched_u(v, ux, uy, I, Gx_freq, Gy_freq, erosion, dx, ws, dLdI=None, dLdE=None,
                      stream=None):
    """Batched adjoint from the u stored by forward_batched_u."""
    torch = _torch()
    B, N = I.shape[0], I.shape[-1]
    W = N - 2 * erosion
    if dLdI is None:
        dLdI = torch.empty((B, N, N), dtype=torch.float32, device=I.device)
    if dLdE is None:
        dLdE = torch.empty((B, N, N), dtype=torch.float32, device=I.device)
    img, val = B * N * N, B * W * W
    call("srkAdjointBatchedU", _dev(v, val, "v"), _dev(ux, val, "ux"), _dev(uy, val, "uy"),
         _dev(I, img, "I"), _spec(Gx_freq, ws, "Gx"), _spec(Gy_freq, ws, "Gy"), N, W, erosion,
         float(dx), B, _stream(stream), _dev(dLdI, img, "dLdI"), _dev(dLdE, img, "dLdE"),
         ws.handle)
    return dLdI, dLdE


def tangent_batched(f, g, df, dg, Gx_freq, Gy_freq, erosion, dx, ws, dh=None, stream=None):
    torch = _torch()
    B, N = f.shape[0], f.shape[-1]
    W = N - 2 * erosion
    if dh is None:
        dh = torch.empty((B, W, W), dtype=torch.float32, device=f.device)
    img, val = B * N * N, B * W * W
    call("srkTangentBatched", _dev(f, img, "f"), _dev(g, img, "g"), _dev(df, img, "df"),
         _dev(dg, img, "dg"), _spec(Gx_freq, ws, "Gx"), _spec(Gy_freq, ws, "Gy"), N, W, erosion,
         float(dx), B, _stream(stream), _dev(dh, val, "dh"), ws.handle)
    return dh


def tangent_adjoint_batched(f, g, df, dg, lam_h, dlam_h, Gx_freq, Gy_freq, erosion, dx, ws,
                            dLam_f=None, dLam_g=None, stream=None):
    torch = _torch()
    B, N = f.shape[0], f.shape[-1]
    W = N - 2 * erosion
    if dLam_f is None:
        dLam_f = torch.empty((B, N, N), dtype=torch.float32, device=f.device)
    if dLam_g is None:
        dLam_g = torch.empty((B, N, N), dtype=torch.float32, device=f.device)
    img, val = B * N * N, B * W * W
    call("srkTangentAdjointBatched", _dev(f, img, "f"), _dev(g, img, "g"), _dev(df, img, "df"),
         _dev(dg, img, "dg"), _dev(lam_h, val, "lam_h"), _dev(dlam_h, val, "dlam_h"),
         _spec(Gx_freq, ws, "Gx"), _spec(Gy_freq, ws, "Gy"), N, W, erosion, float(dx), B,
         _stream(stream), _dev(dLam_f, img, "dLam_f"), _dev(dLam_g, img, "dLam_g"), ws.handle)
    return dLam_f, dLam_g


# ---- critical-dimension measurement ------------------------------------------

FLT_MAX = 3.4028234663852886e38


def cd_edge_found(px, L: int = 16):
    """True where a CD edge (tensor or array of pixels) was found: a missing edge
    reads -FLT_MAX / L or +FLT_MAX / L."""
    return abs(px) < 0.5 * FLT_MAX / L


def build_polyphase_table(L: int = 16, A: int = 8, stream=None):
    """BuildPolyphaseTable: [L, 2A+1] float32 device table, rows summing to 1."""
    torch = _torch()
    out = torch.empty((L, 2 * A + 1), dtype=torch.float32, device="cuda")
    call("srkBuildPolyphaseTable", L, A, _ptr(out), _stream(stream))
    return out


def upsample_batched(sig, coeff, L: int = 16, A: int = 8, stream=None):
    """UpsampleBatched: [B, N] device signals -> [B, N*L] with the table coeff."""
    torch = _torch()
    sig = sig.contiguous()
    B, N = sig.shape
    out = torch.empty((B, N * L), dtype=torch.float32, device=sig.device)
    call("srkUpsampleBatched", _dev(sig, B * N, "signals"), _dev(coeff, L * (2 * A + 1), "coeff"),
         B, N, L, A, _ptr(out), _stream(stream))
    return out


def measure_cd_batched_ex(sig=None, coeff=None, up=None, N=None, L: int = 16, A: int = 8,
                          stream=None):
    """MeasureCD_Batched overload with a prebuilt table and / or upsampled
    signals (zero crossings); sig [B, N], up [B, N*L] (N given or taken from sig)."""
    torch = _torch()
    if sig is not None:
        sig = sig.contiguous()
        B, N = sig.shape
    else:
        B = up.shape[0]
    out = [torch.empty(B, dtype=torch.float32, device="cuda") for _ in range(3)]
    call("srkMeasureCDBatchedEx", _dev(sig, B * N, "signals", optional=True),
         _dev(coeff, L * (2 * A + 1), "coeff", optional=True),
         _dev(up, B * N * L, "upsampled", optional=True), B, N, L, A, *map(_ptr, out),
         _stream(stream))
    return tuple(out)


def measure_cd_batched(sig, L: int = 16, A: int = 8, threshold: float = 0.0, stream=None):
    """MeasureCD_Batched of [B, N] row-major device signals: (left_px, right_px,
    cd_px) device tensors [B]; a missing edge reads -FLT_MAX / L or +FLT_MAX / L."""
    torch = _torch()
    sig = sig.contiguous()
    B, N = sig.shape
    out = [torch.empty(B, dtype=torch.float32, device=sig.device) for _ in range(3)]
    call("srkMeasureCDBatched", _dev(sig, B * N, "signals"), B, N, L, A, float(threshold),
         *map(_ptr, out), _stream(stream))
    return tuple(out)


def shrinkage_cd_batched(E, R, erosion: int, coeff_c: float = 1.0, center_row: bool = True,
                         L: int = 16, A: int = 8, threshold: float = 0.0, stream=None):
    """CD of E_crop + coeff_c R on the centre row (or column) of the W x W valid
    region; E [B, N, N], R [B, W, W] device tensors in the reference layout."""
    torch = _torch()
    B, N = E.shape[0], E.shape[-1]
    W = R.shape[-1]
    out = [torch.empty(B, dtype=torch.float32, device=E.device) for _ in range(3)]
    call("srkShrinkageCDBatched", _dev(E, B * N * N, "E"), _dev(R, B * W * W, "R"), N, W,
         erosion, B, float(coeff_c),
         int(bool(center_row)), L, A, float(threshold), *map(_ptr, out), _stream(stream))
    return tuple(out)


def _make_function():
    torch = _torch()

    class ShrinkageFunction(torch.autograd.Function):
        """Autograd pairing of the forward with its hand-written adjoint.

        h = ShrinkageFunction.apply(f, g, Gx_freq, Gy_freq, erosion, dx, ws):
        the backward recomputes u from g (nothing but f and g is saved), so
        peak memory stays at the inputs and outputs."""

        @staticmethod
        def forward(ctx, f, g, Gx_freq, Gy_freq, erosion, dx, ws):
            f = f.contiguous()
            g = g.contiguous()
            ctx.save_for_backward(f, g)
            ctx.cfg = (Gx_freq, Gy_freq, erosion, dx, ws)
            return forward_batched(g, f, Gx_freq, Gy_freq, erosion, dx, ws)

        @staticmethod
        def backward(ctx, v):
            f, g = ctx.saved_tensors
            Gx_freq, Gy_freq, erosion, dx, ws = ctx.cfg
            lam_f, lam_g = adjoint_batched(v.contiguous(), g, f, Gx_freq, Gy_freq, erosion, dx, ws)
            return lam_f, lam_g, None, None, None, None, None

    return ShrinkageFunction


def __getattr__(name):
    if name == "ShrinkageFunction":
        fn = _make_function()
        globals()["ShrinkageFunction"] = fn
        return fn
    raise AttributeError(name)
