"""Benchmark: SRK shrinkage forward + adjoint throughput on one MI355X.

Workload (one step): srkTrainingForward over B=5000 images of N=512 (E = I,
Gaussian random fields), KW=121 (erosion 60, W=392), FP32, followed by the
batched adjoint (lambda_f, lambda_g) for a random cotangent with u recomputed.
Metric: input Mpix/s = B*N^2 / (t_fwd + t_adj).  A second point, N=16384
single image (f != g), is reported under "secondary".

    python bench.py [--gpus 1] [--steps 5] [--warmup 2]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def pmc_traffic(N, B, kw, M, precision="fp32"):
    """HBM bytes per fwd+adj step from the committed rocprofv3 --pmc summary of
    this workload (FETCH_SIZE and WRITE_SIZE, copy-calibrated), or None."""
    tag = "" if precision == "fp32" else "_" + precision
    path = os.path.join(PROFILES, f"hbm_traffic_fwd_adj_n{N}_b{B}_kw{kw}_m{M}{tag}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_step:"):
            kv = dict(x.split("=") for x in line.split(":", 1)[1].split())
            return int(float(kv["fetch"]) + float(kv["write"]))
    return None


def alg_bytes(N, W, B, e_eq_i):
    """Compulsory fp32 bytes of one forward + adjoint (u recomputed):
    f != g: read f, g (8N^2), write h (4W^2); read f, g, v (8N^2 + 4W^2), write
    lambda_f, lambda_g (8N^2) -> 24N^2 + 8W^2.  E = I: 16N^2 + 8W^2."""
    per = (16 if e_eq_i else 24) * N * N + 8 * W * W
    return per * B


def pass_bytes_per_tile(N, M, e):
    """Compulsory HBM bytes of each pass for one whole-image tile (N = M, E = I,
    FP32 planes): what the pass must read and write given the three-pass
    design (input images, the exchange planes it consumes and produces,
    outputs), the algorithmic figure its per-launch roofline is priced on.
    Planes hold H = M/2 + 1 complex (8-byte) rows per column; u is stored
    and read on the valid columns only, the adjoint sources likewise."""
    if N != M:
        return None
    W, H = N - 2 * e, M // 2 + 1
    img, val = 4 * N * N, 4 * W * W
    x1 = 8 * M * H                 # one X1 plane, all columns
    col = 8 * H                    # one X2 / X1 column
    src_cols = 4 * sum(1 for b in range(M // 4) if 4 * b + 3 >= e and 4 * b < N - e)
    return {
        "k_p1_cols": img + x1,
        "k_p1_adj_src": img + val + 2 * col * src_cols,
        "k_p2_conv": x1 + 2 * col * W,
        "k_p2_corr": 2 * col * W + col * (N + 2),
        "k_p3_fwd": 2 * col * W + 4 * N * (W + 2) + val,
        "k_p3_adj": 2 * col * W + col * (N + 2) + val + 2 * img,
    }


FP32_VECTOR_PEAK_TFS = 157.3   # MI355X FP32 vector (packed); 78.6 without packed FP32 ops


def nominal_flops_per_tile(M):
    """Nominal FLOPs of one tile through one forward and one adjoint (u
    recomputed): 5 n log2 n per complex n-point FFT, real-input transforms
    counted half, plus the spectrum products; fwd = 15 M^2 log2 M + 6 M^2,
    adj = 30 M^2 log2 M + 12 M^2."""
    lg = M.bit_length() - 1
    return 45 * M * M * lg + 18 * M * M


def ref_train_ws_bytes(N, B):
    """Device bytes of the reference's batched forward workspace
    (SrkTrainWorkspace::allocate, training/srk_training_forward.cuh:130-141): five
    real [B x N x N] arrays and two complex [B x N x (N/2+1)] arrays; cuFFT work
    areas are not counted."""
    return 5 * 4 * B * N * N + 2 * 8 * B * N * (N // 2 + 1)


def device_copy_GBps(torch, mb=1024, reps=10):
    """Read + write bandwidth of a plain device-to-device copy of `mb` MiB (the
    rate a bandwidth-bound pass can reach on this device)."""
    n = mb * (1 << 20) // 4
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    s, t = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        b.copy_(a)
    t.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(t) / reps
    del a, b
    return 2 * n * 4 / (ms / 1e3) / 1e9


def cpu_cores():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def cpu_model():
    """Model name of the host CPU from /proc/cpuinfo (platform fallback)."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_share():
    """Host threads this process may use: the CPU share the machine grants one
    GPU job (OMP_NUM_THREADS where it is set), else every core of the affinity
    mask."""
    aff = cpu_cores()
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        share = 0
    return min(aff, share) if share > 0 else aff


def cpu_baseline(N, kw, sigma, e, budget_s=12.0):
    """Timed FP64 FFT-form forward + adjoint of the oracle on the host cores."""
    import numpy as np
    import oracle
    from shrinkage_full_tests_amd import patterns
    threads = cpu_share()
    oracle.set_threads(threads)
    Kx, Ky = oracle.srsrk_kernel(kw, 1.0, sigma)
    W = N - 2 * e
    f = patterns.grf(N, 1000, B=1).astype(np.float64)
    v = np.random.default_rng(5).standard_normal((1, W, W))
    t0 = time.perf_counter()
    oracle.fwd_adj_fft_batch(f, f, v, e, Kx, Ky)
    t1 = time.perf_counter() - t0
    n = max(1, min(400, int(budget_s / max(t1, 1e-3))))
    fb = np.repeat(f, n, axis=0)
    vb = np.repeat(v, n, axis=0)
    t0 = time.perf_counter()
    oracle.fwd_adj_fft_batch(fb, fb, vb, e, Kx, Ky)
    dt = time.perf_counter() - t0
    return {"value": round(n * N * N / dt / 1e6, 3), "unit": "Mpix/s", "cores": threads,
            "kind": "port",
            "cpu_model": cpu_model(),
            "cpu_count": os.cpu_count(),
            "affinity_cores": cpu_cores(),
            "threads_used": threads,
            "sample": f"{n} image(s) N={N} KW={kw} E=I, FP64 FFT-form fwd+adj (oracle), "
                      f"{dt:.1f} s on {threads} threads (the CPU share of one GPU job; "
                      f"{cpu_cores()} cores in the affinity mask, os.cpu_count() = "
                      f"{os.cpu_count()}) of {cpu_model()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--b", type=int, default=5000)
    ap.add_argument("--kw", type=int, default=121)
    ap.add_argument("--m", type=int, default=0, help="tile/FFT size (default N)")
    ap.add_argument("--chunk", type=int, default=0, help="tiles per chunk (0 = auto)")
    ap.add_argument("--p2-rows", type=int, default=0, help="k_y bands per row-pass block (0 = auto)")
    ap.add_argument("--scratch-mb", type=int, default=0)
    ap.add_argument("--concurrency", type=int, default=-1,
                    help="adjoint front passes on two streams (1) or one (0); -1 = library default")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--precision", default="both",
                    choices=("fp32", "mixed_fp16", "bf16_exchange", "both"),
                    help="both = FP32 plus the two reduced-precision forms")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--large-n", type=int, default=16384)
    ap.add_argument("--large-m", type=int, default=0,
                    help="tile size of the large image (0: time 512 and 1024, report the faster)")
    ap.add_argument("--no-ragged", action="store_true",
                    help="skip the ragged 64-instance fwd+adj+tangent-adjoint point")
    ap.add_argument("--f-ne-g", action="store_true",
                    help="main point with distinct images for f (I) and g (E)")
    ap.add_argument("--no-fused", action="store_true",
                    help="skip the fused one-launch pipeline timing (one-tile images, M <= 512)")
    ap.add_argument("--no-fwd-configs", action="store_true",
                    help="skip the forward-only sweep over the batched training shapes")
    ap.add_argument("--no-derivatives", action="store_true",
                    help="skip the single-image derivative operators (N 1024-4096, KW 61)")
    ap.add_argument("--no-stored-u", action="store_true",
                    help="skip the stored-activation (u kept by the forward) fwd+adj point")
    ap.add_argument("--no-cd", action="store_true",
                    help="skip the batched CD measurement on the main point's output")
    ap.add_argument("--no-refcfg", action="store_true",
                    help="skip the forward at the configuration of the reference's published timing")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="run one plain device copy of this size first (PMC traffic calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the shrinkage operator library has no "
              "multi-GPU path (sharding the batch across GPUs is described, not "
              "implemented). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import shrinkage_full_tests_amd as srk
    from shrinkage_full_tests_amd import patterns

    torch.cuda.set_device(0)
    srk.lib()  # fails loudly if the HIP library is missing
    N, B, kw = args.n, args.b, args.kw
    e = (kw - 1) // 2
    W = N - 2 * e
    sigma = e / 3.0
    M = args.m or N
    dx = 1.0

    if args.calib_copy_mb:
        n = args.calib_copy_mb * (1 << 20) // 4
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    gx, gy = srk.kernel_spectra_srsrk(kw, sigma, M)
    E = patterns.grf_torch(B, N, seed=1000)
    # f = I (differentiated) and g = E (convolved): the same images unless --f-ne-g
    I = patterns.grf_torch(B, N, seed=2000) if args.f_ne_g else E
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    v = torch.randn((B, W, W), generator=gen, device="cuda")
    R = torch.empty((B, W, W), device="cuda")
    lf = torch.empty((B, N, N), device="cuda")
    lg = torch.empty((B, N, N), device="cuda")
    ws = srk.Workspace(N, W, e, M, B_max=B, scratch_mb=args.scratch_mb)
    step_tensors = [E, v, R, lf, lg] + ([I] if args.f_ne_g else [])
    memory = {
        "workspace_bytes": ws.nbytes,
        "caller_buffers_bytes": sum(t.numel() * t.element_size() for t in step_tensors),
        "reference_training_workspace_bytes": ref_train_ws_bytes(N, B),
        "note": "device memory the library holds for the whole fwd+adj step (exchange "
                "buffers, spectra, tile tables) against the reference's forward-only batched "
                "workspace at the same N and B; the inputs and outputs are the caller's",
    }
    ws.set_tuning(chunk_tiles=args.chunk, p2_rows=args.p2_rows,
                  concurrency=None if args.concurrency < 0 else args.concurrency)
    stream = torch.cuda.current_stream()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def fwd():
        if M == N:
            srk.srkTrainingForward(E, I, gx, gy, N, W, e, dx, B, ws, R)
        else:
            srk.forward_batched(E, I, gx, gy, e, dx, ws, R)

    def adj():
        srk.adjoint_batched(v, E, I, gx, gy, e, dx, ws, lf, lg)

    def step():
        fwd()
        adj()

    def measure(precision):
        """(ms per step, ms fwd, ms adj, hipgraph used) for one precision form."""
        ws.set_precision(precision)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        graph = None
        if not args.no_graph:
            graph = torch.cuda.CUDAGraph()
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                with torch.cuda.graph(graph, stream=s):
                    step()
            torch.cuda.current_stream().wait_stream(s)
            graph.replay()
            torch.cuda.synchronize()
        run = graph.replay if graph is not None else step
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(args.steps):
            run()
        stop.record(stream)
        torch.cuda.synchronize()
        ms_ = start.elapsed_time(stop) / args.steps
        # per-kernel launch durations: the same steps run eagerly with every
        # launch bracketed by HIP events on the launch stream, on one lane (with
        # two lanes the launches of alternate chunks overlap and each one's
        # duration counts the other's work)
        ws.set_tuning(concurrency=0)
        ws.set_timing(True)
        ws.kernel_times()
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(args.steps):
            step()
        stop.record(stream)
        torch.cuda.synchronize()
        ms_eager = start.elapsed_time(stop) / args.steps
        kt = ws.kernel_times()
        ws.set_timing(False)
        ws.set_tuning(concurrency=1 if args.concurrency < 0 else args.concurrency)
        kernels = {}
        for name, (tot, cnt, tiles) in kt.items():
            if cnt:
                kernels[name] = {"ms_per_step": round(tot / args.steps, 3),
                                 "launches_per_step": cnt // args.steps,
                                 "avg_launch_us": round(1e3 * tot / cnt, 1),
                                 "tiles_per_launch": round(tiles / cnt, 1)}
        # forward / adjoint split (eager, same stream)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        reps = max(1, min(3, args.steps))
        tf = ta = 0.0
        for _ in range(reps):
            ev[0].record(stream)
            fwd()
            ev[1].record(stream)
            adj()
            ev[2].record(stream)
            torch.cuda.synchronize()
            tf += ev[0].elapsed_time(ev[1]) / reps
            ta += ev[1].elapsed_time(ev[2]) / reps
        del graph
        return ms_, tf, ta, ms_eager, kernels

    dtypes = {"mixed_fp16": "f16 MFMA operands, f32 accumulate",
              "bf16_exchange": "f32 transforms, bf16 inter-pass spectra (adjoint u in f32)"}
    reduced = {}
    for prec in ("mixed_fp16", "bf16_exchange"):
        if args.precision not in (prec, "both") or (prec == "mixed_fp16" and M > 1024):
            continue
        rms, rtf, rta, _, rk = measure(prec)
        reduced[prec] = {"value": round(B * N * N / (rms / 1e3) / 1e6, 1), "unit": "Mpix/s",
                         "ms_per_step": round(rms, 3), "ms_fwd": round(rtf, 3),
                         "ms_adj": round(rta, 3), "dtype": dtypes[prec],
                         "tolerance": "tol_rel 1e-2, tol_abs 1e-3 vs FP64 (tests/test_gpu_mixed.py)",
                         "kernels": rk}
    if args.precision in reduced:
        r = reduced[args.precision]
        ms, t_f, t_a, ms_eager, kernels = (r["ms_per_step"], r["ms_fwd"], r["ms_adj"], None,
                                           r["kernels"])
    else:
        ms, t_f, t_a, ms_eager, kernels = measure("fp32")
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"]) if kernels else None
    pbt = pass_bytes_per_tile(N, M, e)
    if pbt:
        for name, kv in kernels.items():
            if name in pbt:
                per_launch = pbt[name] * kv["tiles_per_launch"]
                kv["pass_bytes_per_launch"] = int(per_launch)
                kv["pass_GBps"] = round(per_launch / (kv["avg_launch_us"] * 1e3), 1)
                kv["pass_frac"] = round(kv["pass_GBps"] / HBM_PEAK_GBS, 4)
    for prec, r in reduced.items():
        r["speedup_vs_fp32"] = round(ms / r["ms_per_step"], 3)
        tr = pmc_traffic(N, B, kw, M, prec)
        r["traffic"] = tr
        if tr:
            r["traffic_TB_per_s"] = round(tr / (r["ms_per_step"] / 1e3) / 1e12, 2)
    graph = not args.no_graph

    value = B * N * N / (ms / 1e3) / 1e6
    nbytes = alg_bytes(N, W, B, not args.f_ne_g)
    achieved = nbytes / (ms / 1e3) / 1e9
    info = ws.info()
    traffic = pmc_traffic(N, B, kw, M, "fp32")
    copy_gbs = device_copy_GBps(torch)
    traffic_gbs = traffic / (ms / 1e3) / 1e9 if traffic else None

    stored = None
    if not args.no_stored_u:
        # the reference's stored-activation flow, batched: the forward keeps
        # u_x, u_y (8 W^2 bytes per image) and the adjoint reads them
        ws.set_precision("fp32")
        ux = torch.empty((B, W, W), device="cuda")
        uy = torch.empty((B, W, W), device="cuda")

        def step_u():
            srk.forward_batched_u(E, I, gx, gy, e, dx, ws, R, ux, uy)
            srk.adjoint_batched_u(v, ux, uy, I, gx, gy, e, dx, ws, lf, lg)

        for _ in range(args.warmup):
            step_u()
        torch.cuda.synchronize()
        gu = None
        if not args.no_graph:
            gu = torch.cuda.CUDAGraph()
            su = torch.cuda.Stream()
            su.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(su):
                with torch.cuda.graph(gu, stream=su):
                    step_u()
            torch.cuda.current_stream().wait_stream(su)
            gu.replay()
            torch.cuda.synchronize()
        run_u = gu.replay if gu is not None else step_u
        start.record(stream)
        for _ in range(args.steps):
            run_u()
        stop.record(stream)
        torch.cuda.synchronize()
        msu = start.elapsed_time(stop) / args.steps
        stored = {"workload": "srkForwardBatchedU + srkAdjointBatchedU (u stored by the "
                              "forward, as the reference's srkForwardOaS + adjointSRK), FP32",
                  "value": round(B * N * N / (msu / 1e3) / 1e6, 1), "unit": "Mpix/s",
                  "ms_per_step": round(msu, 3),
                  "speedup_vs_recompute": round(ms / msu, 3),
                  "stored_activation_bytes": 8 * B * W * W}
        del gu, ux, uy

    fwd_cfg = None
    if not args.no_fwd_configs and (N, B, kw) == (512, 5000, 121):
        # forward-only Mpix/s over the batched training shapes (B = 5000,
        # N in {256, 512}, KW in {31, 61, 121}, E = I; dx changes no work),
        # srkTrainingForward replayed from a hipGraph
        fwd_cfg = {"workload": "srkTrainingForward, B=5000, E=I, FP32, one hipGraph replay "
                               "per call", "points": []}
        ws.set_precision("fp32")
        for Nf in (256, 512):
            Ef = E if Nf == N else patterns.grf_torch(B, Nf, seed=1100)
            for kwf in (31, 61, 121):
                ef = (kwf - 1) // 2
                Wf = Nf - 2 * ef
                gxf, gyf = srk.kernel_spectra_srsrk(kwf, ef / 3.0, Nf)
                wsf = srk.Workspace(Nf, Wf, ef, Nf, B_max=B)
                Rf = torch.empty((B, Wf, Wf), device="cuda")

                def ff():
                    srk.srkTrainingForward(Ef, Ef, gxf, gyf, Nf, Wf, ef, dx, B, wsf, Rf)

                ff()
                torch.cuda.synchronize()
                gf = torch.cuda.CUDAGraph()
                sf = torch.cuda.Stream()
                sf.wait_stream(torch.cuda.current_stream())
                with torch.cuda.stream(sf):
                    with torch.cuda.graph(gf, stream=sf):
                        ff()
                torch.cuda.current_stream().wait_stream(sf)
                gf.replay()
                torch.cuda.synchronize()
                start.record(stream)
                for _ in range(5):
                    gf.replay()
                stop.record(stream)
                torch.cuda.synchronize()
                msf_ = start.elapsed_time(stop) / 5
                bf = (4 * Nf * Nf + 4 * Wf * Wf) * B   # E = I: read the image once, write R
                fwd_cfg["points"].append({
                    "N": Nf, "KW": kwf, "W": Wf, "ms_per_call": round(msf_, 3),
                    "Mpix_per_s": round(B * Nf * Nf / (msf_ / 1e3) / 1e6, 1),
                    "valid_Mpix_per_s": round(B * Wf * Wf / (msf_ / 1e3) / 1e6, 1),
                    "hbm_frac": round(bf / (msf_ / 1e3) / 1e9 / HBM_PEAK_GBS, 4)})
                del gf, wsf, Rf
            del Ef
        torch.cuda.empty_cache()

    fused_out = None
    if not args.no_fused and M == N and M in (256, 512):
        # the opt-in fused one-launch pipeline (SRK_TUNE_FUSED) on the same step, eager
        ws.set_precision("fp32")
        ws.set_tuning(fused=1)
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(args.steps):
            step()
        stop.record(stream)
        torch.cuda.synchronize()
        msf = start.elapsed_time(stop) / args.steps
        available, timeouts = ws.fused_status()
        ws.set_tuning(fused=0)
        fused_out = {"workload": "the same fwd+adj step with SRK_TUNE_FUSED = 1 (one persistent "
                                 "launch per operator, exchange ring of 4 slots per XCD, "
                                 "stage lag 2), FP32, eager",
                     "ms_per_step": round(msf, 3), "speedup_vs_three_pass": round(ms / msf, 3),
                     "timed_out_waits": timeouts}

    cd_out = None
    if not args.no_cd:
        # the training CD sweep's measurement (training/main_cd_training.cu:213-261):
        # crop, E + c R, centre row, polyphase upsampling, edges and CD of every image,
        # fused into one launch over the batch
        def cd_call():
            return srk.shrinkage_cd_batched(E, R, e, 1.0, center_row=True, threshold=0.5)

        cd_call()
        torch.cuda.synchronize()
        reps = 20
        start.record(stream)
        for _ in range(reps):
            cd_call()
        stop.record(stream)
        torch.cuda.synchronize()
        ms_cd = start.elapsed_time(stop) / reps
        cd_out = {"workload": f"srkShrinkageCDBatched: centre-row CD of E_crop + c R, B={B}, "
                              f"W={W}, L=16, A=8, threshold 0.5",
                  "ms_per_call": round(ms_cd, 4),
                  "signals_per_s": round(B / (ms_cd / 1e3), 1)}

    secondary = None
    if not args.no_large:
        del lf, lg, R, v, E, ws
        torch.cuda.empty_cache()
        NL = args.large_n
        WL = NL - 2 * e
        # tile sizes timed for the large image (the faster one is reported per
        # form, as the autotuner would pick it); --large-m fixes one
        tiles_L = [args.large_m] if args.large_m else [512, 1024]
        fL = patterns.grf_torch(1, NL, seed=3000, sigma=8.0)
        gL = patterns.grf_torch(1, NL, seed=4000, sigma=8.0)
        vL = torch.randn((1, WL, WL), generator=gen, device="cuda")
        hL = torch.empty((1, WL, WL), device="cuda")
        lfL = torch.empty((1, NL, NL), device="cuda")
        lgL = torch.empty((1, NL, NL), device="cuda")
        specL = {ML: srk.kernel_spectra_srsrk(kw, sigma, ML) for ML in tiles_L}
        wsL = {ML: srk.Workspace(NL, WL, e, ML, B_max=1) for ML in tiles_L}

        eager_ms = {}  # (precision, M) -> eager ms per step

        def time_large(precision, ML):
            w_, (gxl, gyl) = wsL[ML], specL[ML]
            w_.set_precision(precision)

            def stepL():
                srk.forward_batched(gL, fL, gxl, gyl, e, dx, w_, hL)
                srk.adjoint_batched(vL, gL, fL, gxl, gyl, e, dx, w_, lfL, lgL)

            # timed as the main point: a hipGraph of the step, replayed (the
            # eager time is kept beside it in eager_ms)
            stepL()
            torch.cuda.synchronize()
            start.record(stream)
            for _ in range(3):
                stepL()
            stop.record(stream)
            torch.cuda.synchronize()
            eager_ms[(precision, ML)] = start.elapsed_time(stop) / 3
            if args.no_graph:
                return eager_ms[(precision, ML)]
            gr = torch.cuda.CUDAGraph()
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                with torch.cuda.graph(gr, stream=s):
                    stepL()
            torch.cuda.current_stream().wait_stream(s)
            gr.replay()
            torch.cuda.synchronize()
            reps = max(3, args.steps)
            start.record(stream)
            for _ in range(reps):
                gr.replay()
            stop.record(stream)
            torch.cuda.synchronize()
            del gr
            return start.elapsed_time(stop) / reps

        def best_large(precision):
            by_m = {ML: time_large(precision, ML) for ML in tiles_L
                    if not (precision == "mixed_fp16" and ML > 1024)}
            if not by_m:
                return None, None, {}
            ML = min(by_m, key=by_m.get)
            return by_m[ML], ML, {str(m): round(t, 3) for m, t in by_m.items()}

        bL = alg_bytes(NL, WL, 1, False)
        msL, MLb, byL = best_large("fp32")
        secondary = {"workload": f"fwd+adj N={NL} B=1 KW={kw} f!=g",
                     "value": round(NL * NL / (msL / 1e3) / 1e6, 1), "unit": "Mpix/s",
                     "ms_per_step": round(msL, 3), "M": MLb, "ms_by_M": byL,
                     "ms_per_step_eager": round(eager_ms[("fp32", MLb)], 3),
                     "timing": "hipGraph replay" if not args.no_graph else "eager",
                     "hbm_frac": round(bL / (msL / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
        for prec in reduced:
            msM, MLm, byM = best_large(prec)
            if msM is None:
                continue
            secondary[prec] = {"value": round(NL * NL / (msM / 1e3) / 1e6, 1),
                               "ms_per_step": round(msM, 3), "M": MLm, "ms_by_M": byM,
                               "ms_per_step_eager": round(eager_ms[(prec, MLm)], 3),
                               "hbm_frac": round(bL / (msM / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                               "speedup_vs_fp32": round(msL / msM, 3)}

    deriv = None
    if not args.no_derivatives and (N, B, kw) == (512, 5000, 121):
        # the differentiated single-image path of the derivative checks
        # (tan_tanadj/main_tan_tanadj.cu): forward, adjoint (u recomputed),
        # tangent and tangent-adjoint, B = 1, KW = 61, f != g, M = 1024 tiles
        if not args.no_large:
            del fL, gL, vL, hL, lfL, lgL, wsL
        torch.cuda.empty_cache()
        deriv = {"workload": "B=1, KW=61 (erosion 30), f != g GRF, M=1024, FP32, eager calls",
                 "points": []}
        kwd = 61
        ed = (kwd - 1) // 2
        gxd, gyd = srk.kernel_spectra_srsrk(kwd, ed / 3.0, 1024)
        for Nd in (1024, 2048, 4096):
            Wd = Nd - 2 * ed
            f1, g1, df1, dg1 = (patterns.grf_torch(1, Nd, seed=7000 + 10 * Nd + i) for i in range(4))
            vd = torch.randn((1, Wd, Wd), generator=gen, device="cuda")
            wd = torch.randn((1, Wd, Wd), generator=gen, device="cuda")
            wsd = srk.Workspace(Nd, Wd, ed, 1024, B_max=1)
            ops = {
                "forward": lambda: srk.forward_batched(g1, f1, gxd, gyd, ed, dx, wsd),
                "adjoint": lambda: srk.adjoint_batched(vd, g1, f1, gxd, gyd, ed, dx, wsd),
                "tangent": lambda: srk.tangent_batched(f1, g1, df1, dg1, gxd, gyd, ed, dx, wsd),
                "tangent_adjoint": lambda: srk.tangent_adjoint_batched(
                    f1, g1, df1, dg1, vd, wd, gxd, gyd, ed, dx, wsd),
            }
            pt = {"N": Nd}
            for name, op in ops.items():
                op()
                torch.cuda.synchronize()
                start.record(stream)
                for _ in range(5):
                    op()
                stop.record(stream)
                torch.cuda.synchronize()
                msd = start.elapsed_time(stop) / 5
                pt[name] = {"ms": round(msd, 3), "Mpix_per_s": round(Nd * Nd / (msd / 1e3) / 1e6, 1)}
            deriv["points"].append(pt)
            wsd.close()
            del f1, g1, df1, dg1, vd, wd
        torch.cuda.empty_cache()

    ragged_out = None
    if not args.no_ragged:
        if not args.no_large and deriv is None:
            del fL, gL, vL, hL, lfL, lgL, wsL
        torch.cuda.empty_cache()
        from shrinkage_full_tests_amd import ragged
        rb = ragged.RaggedBatch()
        ragged_out = {"workload": "64 instances, N in {1024, 2048, 4096}, KW in {31, 61, 91, 121} "
                                  "(seed 4242): fwd + adjoint + tangent-adjoint, one hipGraph",
                      "pixels": rb.pixels}
        for prec in ("fp32",) + tuple(reduced):
            rb.set_precision(prec)
            gr = rb.capture()
            torch.cuda.synchronize()
            start.record(stream)
            for _ in range(3):
                gr.replay()
            stop.record(stream)
            torch.cuda.synchronize()
            msR = start.elapsed_time(stop) / 3
            ragged_out[prec] = {"ms_per_step": round(msR, 3),
                                "Mpix_per_s": round(rb.pixels / (msR / 1e3) / 1e6, 1)}
            del gr
        for prec in reduced:
            ragged_out[prec]["speedup_vs_fp32"] = round(
                ragged_out["fp32"]["ms_per_step"] / ragged_out[prec]["ms_per_step"], 3)
        del rb

    refcfg = None
    if not args.no_refcfg:
        # The reference's one published timing: srkTrainingForward, B=3000, N=256,
        # KW=121 (sigma 20, erosion 60, W=136), dx=1, forward only, 55.39 ms per
        # call on an sm_86 GPU (notebooks/plot_compare_demo.ipynb:20-21).
        Nr, Br, kwr, er = 256, 3000, 121, 60
        Wr = Nr - 2 * er
        gxr, gyr = srk.kernel_spectra_srsrk(kwr, 20.0, Nr)
        Er = patterns.grf_torch(Br, Nr, seed=5000)
        Ir = patterns.grf_torch(Br, Nr, seed=6000)
        Rr = torch.empty((Br, Wr, Wr), device="cuda")
        wsr = srk.Workspace(Nr, Wr, er, Nr, B_max=Br)

        def fref():
            srk.srkTrainingForward(Er, Ir, gxr, gyr, Nr, Wr, er, 1.0, Br, wsr, Rr)

        for _ in range(2):
            fref()
        torch.cuda.synchronize()
        start.record(stream)
        for _ in range(10):
            fref()
        stop.record(stream)
        torch.cuda.synchronize()
        msr = start.elapsed_time(stop) / 10
        refcfg = {"workload": "srkTrainingForward B=3000 N=256 KW=121 erosion=60 W=136 dx=1, "
                              "FP32, forward only (eager calls)",
                  "ms_per_call": round(msr, 3),
                  "Mpix_per_s": round(Br * Nr * Nr / (msr / 1e3) / 1e6, 1),
                  "reference_ms_per_call": 55.39,
                  "reference_hardware": "NVIDIA sm_86 GPU (published, different hardware)",
                  "ratio_reference_over_this": round(55.39 / msr, 2)}
        wsr.close()
        del Er, Ir, Rr

    cpu = None if args.no_cpu else cpu_baseline(N, kw, sigma, e)

    out = {
        # value: the N=512 B=5000 point; the N=16384 point and the mixed-precision
        # ratio are under "secondary" and "mixed_fp16"
        "metric": ("fwd+adjoint Mpix/s, N=512 B=5000 & N=16k, KW=121; MFMA-mixed vs FP32 at tol"
                   if (N, B, kw) == (512, 5000, 121) else
                   f"fwd+adjoint Mpix/s, N={N} B={B} KW={kw}"),
        "value": round(value, 1),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian random fields, random cotangent)",
        "config": {"workload": f"srkTrainingForward + srkAdjointBatched (u recomputed), "
                               f"N={N} B={B} KW={kw} erosion={e} W={W} "
                               f"{'f != g' if args.f_ne_g else 'E=I'}, FP32",
                   "N": N, "B": B, "KW": kw, "erosion": e, "W": W, "M": M,
                   "tiles_per_chunk_fwd": info["chunk_fwd"], "tiles_per_chunk_adj": info["chunk_adj"],
                   "hipgraph": graph,
                   "precision": args.precision if args.precision in reduced else "fp32"},
        "valid_Mpix_per_s": round(B * W * W / (ms / 1e3) / 1e6, 1),
        "flops": ({"convention": "nominal: 5 n log2 n per complex n-point FFT, real-input "
                                 "transforms half, spectrum products; fwd+adj per tile "
                                 "45 M^2 log2 M + 18 M^2",
                   "TFLOP_per_step": round(nominal_flops_per_tile(M) * B / 1e12, 3),
                   "achieved_TFLOP_per_s": round(nominal_flops_per_tile(M) * B / (ms / 1e3)
                                                 / 1e12, 2),
                   "fp32_vector_peak_TFLOP_per_s": FP32_VECTOR_PEAK_TFS,
                   "frac": round(nominal_flops_per_tile(M) * B / (ms / 1e3) / 1e12
                                 / FP32_VECTOR_PEAK_TFS, 4)}
                  if M == N else None),
        "ms_fwd": round(t_f, 3),
        "ms_adj": round(t_a, 3),
        "memory": memory,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_GBps": round(traffic_gbs, 1) if traffic_gbs else None,
                     "traffic_frac": round(traffic_gbs / HBM_PEAK_GBS, 4) if traffic_gbs else None,
                     "copy_GBps": round(copy_gbs, 1),
                     "traffic_frac_of_copy": round(traffic_gbs / copy_gbs, 4) if traffic_gbs else None,
                     "scope": "one fwd+adj step over the batch (one hipGraph launch of all "
                              "passes); algorithmic bytes 16N^2+8W^2 per image = "
                              f"{nbytes / 1e9:.2f} GB per step; traffic = HBM bytes per step "
                              "from the rocprofv3 FETCH_SIZE/WRITE_SIZE summary in profiles/, "
                              "traffic_GBps = traffic / step time; copy_GBps = a plain 1 GiB "
                              "device copy (read + write) timed in the same run",
                     "dominant_kernel": dom,
                     "dominant_kernel_pass": (
                         {"kernel": dom, "bytes_per_launch": kernels[dom]["pass_bytes_per_launch"],
                          "avg_launch_us": kernels[dom]["avg_launch_us"],
                          "achieved": kernels[dom]["pass_GBps"], "unit": "GB/s",
                          "frac": kernels[dom]["pass_frac"],
                          "note": "the pass's own compulsory bytes (its input planes and "
                                  "images, its output planes) per launch / its average "
                                  "launch duration, launches timed on one lane"}
                         if dom and "pass_GBps" in kernels.get(dom, {}) else None)},
        "ms_per_step_eager_timed_one_lane": round(ms_eager, 3) if ms_eager else None,
        "kernels": kernels,
        "mixed_fp16": reduced.get("mixed_fp16"),
        "bf16_exchange": reduced.get("bf16_exchange"),
        "secondary": secondary,
        "ragged": ragged_out,
        "stored_u": stored,
        "fused_pipeline": fused_out,
        "forward_configs": fwd_cfg,
        "derivatives": deriv,
        "cd_metric": cd_out,
        "reference_config_forward": refcfg,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
