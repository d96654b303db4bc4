/*
 * srk.h -- C ABI of the MI355X-native SRK shrinkage operator library
 * (libsrk_amd.so, gfx950).
 *
 *   h = s * ( conv(Gx, g) * df/dx + conv(Gy, g) * df/dy ),   s = dx^2,
 *
 * evaluated on the valid region W = N - 2*erosion of column-major N x N
 * images (element (row r, col c) at c*N + r; batches outermost).
 *
 * The four drop-in entry points keep the reference's argument lists, with the
 * reference's C++ template parameters (Arch, M, EPT, FPB, FPB_COL) and
 * workspace structs folded into one opaque workspace created for a tile size M.
 * Spectra are in the reference's layout: Ghat[k_x * (M/2+1) + k_y], complex
 * float (re, im), unscaled, as produced by srkPrecomputeKernelFreq.
 *
 * Semantics that differ from the reference, on purpose:
 *   - Calls are asynchronous on `stream` and never allocate or synchronise, so
 *     they can be captured into a hipGraph; call hipStreamSynchronize (or
 *     srkSynchronize) to get the reference's blocking behaviour.
 *   - Every function returns an int status (SRK_OK = 0, negative on error)
 *     instead of void.
 *   - lambda_g (dLdE / dLam_g) is the true adjoint of the forward convolution
 *     (a correlation, product with conj(Ghat)), i.e. the reference CPU brute
 *     force adjoint/main_adjoint.cu:206-246.  The reference GPU adjointWrtE
 *     (adjoint/srk_adjoint.cuh:301-382) multiplies by the un-conjugated
 *     spectrum, which equals this result for even kernels and its negative for
 *     the odd SRSRK kernels.
 *   - The workspace may be shared by all entry points and batch sizes up to
 *     B_max (the reference asserts B == B_max, training/srk_training_forward.cuh:186).
 *   - Calls on one workspace share its buffers.  A call on a stream other than
 *     the previous call's waits for that call (a per-workspace completion
 *     event), so one host thread may use one workspace from several streams.
 *     Calls being captured into a graph are not ordered this way: replaying a
 *     graph while calling on the same workspace from another stream races.
 *     Calls from several host threads on one workspace need the caller's
 *     locking (the reference's workspaces are single-threaded too).
 */
#ifndef SRK_AMD_H
#define SRK_AMD_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
    SRK_OK = 0,
    SRK_ERR_INVALID_ARG = -1,      /* negative / inconsistent scalar arguments */
    SRK_ERR_BAD_DIMS = -2,         /* W != N - 2*erosion, W <= 0, erosion < 1   */
    SRK_ERR_NULL_PTR = -3,         /* a required pointer is NULL                */
    SRK_ERR_MISALIGNED = -4,       /* pointer not aligned to its element type   */
    SRK_ERR_UNSUPPORTED_SIZE = -5, /* tile/FFT size M not in {256,512,1024,2048},
                                      or a CD signal beyond the kernel's limits */
    SRK_ERR_WORKSPACE = -6,        /* workspace made for other N / W / erosion  */
    SRK_ERR_BATCH = -7,            /* B < 1 or B > workspace B_max              */
    SRK_ERR_HIP = -8,              /* HIP runtime error (see srkLastHipError)   */
    SRK_ERR_KERNEL = -9            /* kernel parameters (KW, sigma, dx) invalid */
};

/* precision of the tile transforms */
enum {
    SRK_PREC_FP32 = 0,        /* FP32 radix-4/8/16 transforms on the vector ALUs     */
    SRK_PREC_MIXED_FP16 = 1,  /* DFT stages as FP16 MFMA products, FP32 accumulation,
                                 FP32 twiddles, per-line power-of-two scaling; M <= 1024 */
    SRK_PREC_BF16_EXCHANGE = 2 /* FP32 transforms; the spectra exchanged between passes
                                 through HBM stored as bfloat16 (half the bytes), except
                                 the adjoints' recomputed convolution u, which stays FP32
                                 (lambda_f multiplies it by the cotangent).  Meets the
                                 fp16 tolerance (rel 1e-2, abs 1e-3) on inputs of O(1)
                                 magnitude, not the fp32 one */
};

typedef struct srk_workspace_s* srk_workspace_t;

const char* srkGetErrorString(int status);
int srkLastHipError(void);
int srkGetVersion(void);

/* ---- kernel spectra ---------------------------------------------------- */

/*
 * Replaces srkPrecomputeKernelFreq (srk_forward.cuh:154-203): zero-pad the
 * kw x kh col-major kernel h_G (host, float) into M x M with its centre at
 * offset 0 (the reference's pad + fftshift) and write its 2-D DFT (forward,
 * unnormalised) to d_G_freq[M*(M/2+1)] (device, reference layout).  The DFT is
 * evaluated in double precision on the host.  Synchronous.
 */
int srkPrecomputeKernelFreq(const float* h_G, int kw, int kh, int M, void* d_G_freq,
                            hipStream_t stream);

/*
 * SRSRK Green's tensor spectra built entirely in double precision and cached
 * by (kw, sigma, dx, cw, M): Gx = (g/2pi)(x/r) K1(g r) w(r), Gy likewise, with
 * g = 1/sigma, Blackman radius Rw = cw*sigma (training/main_cd_training.cu:40-77,
 * tan_tanadj/main_tan_tanadj.cu:50-116), written in the reference spectrum
 * layout.  h_Gx / h_Gy (optional, kw*kw floats, col-major) receive the spatial
 * kernels.  Synchronous.
 */
int srkKernelSpectraSRSRK(int kw, double sigma, double dx, double cw, int M, void* d_Gx_freq,
                          void* d_Gy_freq, float* h_Gx, float* h_Gy, hipStream_t stream);

/*
 * Device counterpart of srkKernelSpectraSRSRK (replaces the host round trips of
 * srkPrecomputeKernelFreq, srk_forward.cuh:154-203): the kernel values (FP64
 * K1, Blackman window) and the separable 2-D DFT are all evaluated in FP64
 * on the device; only the spectra (and the optional spatial kernels d_Kx /
 * d_Ky, kw*kw floats, col-major, device) are rounded to FP32.  Asynchronous
 * on the stream, no host transfers; takes stream-ordered scratch
 * (hipMallocAsync / hipFreeAsync) of 2 kw^2 + 2 M + 4 kw (M/2+1) doubles.
 * Same arguments and errors as srkKernelSpectraSRSRK.
 */
int srkKernelSpectraSRSRKDevice(int kw, double sigma, double dx, double cw, int M,
                                void* d_Gx_freq, void* d_Gy_freq, float* d_Kx, float* d_Ky,
                                hipStream_t stream);

/*
 * Support check of a spectrum pair of size Ns (reference layout, device):
 * the relative RMS of the Ns-periodic spatial kernel outside the erosion box
 * |x|, |y| <= erosion, evaluated in FP64 on the device (inverse DFT on the
 * box, Parseval for the whole kernel), is written to *ratio (optional);
 * returns SRK_ERR_KERNEL if it exceeds 1e-4, i.e. if the erosion does not
 * cover the kernel radius (erosion < (KW-1)/2) and the valid region would see
 * wrapped or truncated taps.  erosion must be the workspace's; Ns <= max(N, M)
 * of the workspace.  Synchronous (waits for the stream).  No counterpart in
 * the reference, whose drivers choose erosion >= (KW-1)/2 themselves
 * (training/main_cd_training.cu:302, adjoint/main_adjoint.cu:254).
 */
int srkCheckSpectraSupport(const void* Gx_freq, const void* Gy_freq, int Ns, int erosion,
                           hipStream_t stream, srk_workspace_t ws, double* ratio);

/* ---- workspace ---------------------------------------------------------- */

/*
 * Replaces SrkConvWorkspace / SrkAdjWorkspace / SrkTanWorkspace /
 * SrkTanAdjWorkspace ::allocate (srk_forward.cuh:287-345,
 * adjoint/srk_adjoint.cuh:232-294, tan_tanadj/srk_tangent.cuh:61-85,
 * tan_tanadj/srk_tangent_adjoint.cuh:50-79) and SrkTrainWorkspace::allocate
 * (training/srk_training_forward.cuh:117-158).  One workspace serves every
 * entry point for images of size N with the given erosion, batches up to
 * B_max, FFT/tile size M (the size of the spectra passed to the calls; M = N
 * for srkTrainingForward).  scratch_mb bounds the exchange buffers (0 = 4096, or up to 16384
 * when the tiles of one image need more; allocations never exceed what B_max images need).
 */
int srkWorkspaceCreate(int N, int W, int erosion, int M, int B_max, size_t scratch_mb,
                       srk_workspace_t* ws);
int srkWorkspaceDestroy(srk_workspace_t ws);   /* replaces SrkTrainWorkspace::release */
size_t srkWorkspaceBytes(srk_workspace_t ws);  /* device bytes held */
int srkWorkspaceSetChunk(srk_workspace_t ws, int tiles_per_chunk); /* 0 = automatic */
/* tuning knobs (set by the autotuner; 0 restores the default) */
enum {
    SRK_TUNE_CHUNK_TILES = 0,  /* tiles per launch chunk (exchange buffers bound it)   */
    SRK_TUNE_P2_ROWS = 1,      /* k_y bands per row-pass block: 16, 8 or 4 (M <= 512) */
    SRK_TUNE_CONCURRENCY = 2,  /* 1 (default): chunks alternate between two streams; 0: one */
    SRK_TUNE_FUSED = 3,        /* 1: FP32 forward / adjoint of one-tile images (N <= M,
                                  M = 256 or 512) run as one fused persistent launch;
                                  0 (default): three passes per operator */
    SRK_TUNE_FUSED_SLOTS = 4,  /* fused launch: exchange-ring slots per XCD, in
                                  [lag + 1, 4] (default 4) */
    SRK_TUNE_FUSED_LAG = 5     /* fused launch: rounds between the stages of one image,
                                  in [1, slots - 1] (default 2) */
};
int srkWorkspaceSetTuning(srk_workspace_t ws, int key, int value);
/* SRK_PREC_FP32 (default), SRK_PREC_MIXED_FP16 or SRK_PREC_BF16_EXCHANGE for every
   later call on ws */
int srkWorkspaceSetPrecision(srk_workspace_t ws, int precision);
int srkWorkspaceInfo(srk_workspace_t ws, int* M, int* fwd_tiles_per_image,
                     int* adj_tiles_per_image, int* chunk_fwd, int* chunk_adj);

/*
 * Per-launch timing of the engine kernels (instrumentation; no counterpart in
 * the reference).  With timing enabled, every kernel launched by a later call
 * on ws is bracketed by HIP events on that call's stream (skipped while the
 * stream is being captured into a graph).  srkWorkspaceKernelTimes waits for
 * the recorded launches and writes, per kernel kind k < SRK_KERNEL_KINDS,
 * the summed duration in ms, the number of launches and the tiles they
 * processed (any array may be null), then forgets the records.
 */
#define SRK_KERNEL_KINDS 7
int srkWorkspaceSetTiming(srk_workspace_t ws, int enable);
int srkWorkspaceKernelTimes(srk_workspace_t ws, double* total_ms, int* launches,
                            long long* tiles);
const char* srkKernelKindName(int kind);  /* "k_p1_cols", ..., "k_p3_adj", "k_fused", or "" */

/*
 * Fused pipeline status (instrumentation; no counterpart in the reference):
 * *available = 1 if the workspace can run the fused pipeline (N <= M,
 * M = 256 or 512); *timeouts = waits of the last fused launch that passed
 * their deadline (0 before the first fused launch; otherwise the launch's
 * results are invalid, and SRK_ERR_HIP is returned).  Synchronises the device.
 */
int srkWorkspaceFusedStatus(srk_workspace_t ws, int* available, int* timeouts);

int srkSynchronize(hipStream_t stream);

/* ---- drop-in entry points ---------------------------------------------- */

/*
 * Replaces srkTrainingForward (training/srk_training_forward.cuh:175-241):
 * R_b = dx^2 (conv(E_b,Gx) dI_b/dx + conv(E_b,Gy) dI_b/dy) on the valid region.
 * E, I: [B x N x N]; R: [B x W x W]; spectra at size N (any N, as the
 * reference plans its transforms at any N, training/srk_training_forward.cuh:
 * 130-158).  E may equal I.  This is also the op_test() contract of
 * training/main_cd_training.cu:101-121.
 * With a workspace of tile size M == N the call is asynchronous.  With
 * M != N (e.g. N = 192 or 288 on M = 256 or 512 tiles) the spatial kernel is
 * recovered on the erosion box |x|, |y| <= erosion from the N-size spectra
 * (FP64, on the device) and transformed again at size M; the call then waits
 * for the support measure of srkCheckSpectraSupport and returns
 * SRK_ERR_KERNEL if the kernel reaches beyond the box (not checked while the
 * stream is being captured).
 */
int srkTrainingForward(const float* E, const float* I, const void* Gx_freq, const void* Gy_freq,
                       int N, int W, int erosion, float dx, int B, hipStream_t stream, float* R,
                       srk_workspace_t ws);

/*
 * Replaces srkForwardOaS (srk_forward.cuh:468-501): single-image forward with
 * E = g (convolved) and I = f (differentiated).  ux, uy (optional, [W x W])
 * receive s*conv(E,Gx), s*conv(E,Gy) on the valid region when non-NULL.
 */
int srkForwardOaS(const float* E, const float* I, const void* Gx_freq, const void* Gy_freq,
                  int N, int W, int erosion, float dx, hipStream_t stream, float* R,
                  srk_workspace_t ws, float* ux, float* uy);

/*
 * No counterpart in the reference.  srkForwardOaS for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): no workspace; tile size M in {256, 512,
 * 1024} is passed explicitly with spectra at size M; FP32 transforms.  The
 * tile table, twiddles, imported spectra and exchange buffers live in a
 * static 64 MiB device array that each GPU the library runs on holds until
 * the process ends; tiles stream through it a chunk at a time.  A call that
 * needs more than the array (one tile's exchange planes plus the tables)
 * fails with SRK_ERR_UNSUPPORTED_SIZE.  Allocates nothing, synchronises
 * nothing.  Calls that overlap on one GPU share the array and race.
 */
int srkForwardOaSNoAlloc(const float* E, const float* I, const void* Gx_freq, const void* Gy_freq,
                         int N, int W, int erosion, float dx, int M, hipStream_t stream, float* R,
                         float* ux, float* uy);

/*
 * Replaces adjointSRK (adjoint/srk_adjoint.cuh:402-425): for cotangent v
 * [W x W], dLdI = lambda_f [N x N] from the given ux, uy [W x W]
 * (adjoint/srk_adjoint.cuh:25-75) and dLdE = lambda_g [N x N], the true
 * adjoint with respect to E (see the header comment).
 */
int adjointSRK(const float* v, const float* ux, const float* uy, const float* I,
               const void* Gx_freq, const void* Gy_freq, int N, int W, int erosion, float dx,
               hipStream_t stream, float* dLdI, float* dLdE, srk_workspace_t ws);

/*
 * Replaces shrinkageTangentOaS (tan_tanadj/srk_tangent.cuh:106-151):
 * dh = s( conv(Gx,dg) f_x + conv(Gy,dg) f_y + conv(Gx,g) df_x + conv(Gy,g) df_y ).
 */
int shrinkageTangentOaS(const float* f, const float* g, const float* df, const float* dg,
                        const void* Gx_freq, const void* Gy_freq, int N, int W, int erosion,
                        float dx, hipStream_t stream, float* dh, srk_workspace_t ws);

/*
 * Replaces shrinkageTangentAdjointOaS (tan_tanadj/srk_tangent_adjoint.cuh:100-159):
 * dLam_f = A_f(dlam_h; u(g)) + A_f(lam_h; u(dg)),
 * dLam_g = A_g(dlam_h; f) + A_g(lam_h; df)   (A_g the true adjoint).
 */
int shrinkageTangentAdjointOaS(const float* f, const float* g, const float* df,
                               const float* dg, const float* lam_h, const float* dlam_h,
                               const void* Gx_freq, const void* Gy_freq, int N, int W,
                               int erosion, float dx, hipStream_t stream, float* dLam_f,
                               float* dLam_g, srk_workspace_t ws);

/* ---- batched entry points (no reference counterpart) ------------------- */

/*
 * Batched forward with spectra at the workspace tile size M (any N; N > M
 * runs overlap-and-save tiles).  g = E (convolved), f = I (differentiated).
 */
int srkForwardBatched(const float* E, const float* I, const void* Gx_freq, const void* Gy_freq,
                      int N, int W, int erosion, float dx, int B, hipStream_t stream, float* R,
                      srk_workspace_t ws);

/*
 * Batched adjoint with recomputed activations: (dLdI, dLdE) = (lambda_f,
 * lambda_g) for cotangents v [B x W x W]; u = s*conv(E, G) is recomputed
 * from E instead of being stored by the forward.  The adjoint counterpart of
 * srkTrainingForward; E may equal I.
 */
int srkAdjointBatched(const float* v, const float* E, const float* I, const void* Gx_freq,
                      const void* Gy_freq, int N, int W, int erosion, float dx, int B,
                      hipStream_t stream, float* dLdI, float* dLdE, srk_workspace_t ws);

/*
 * The batched pair of the reference's stored-activation flow (srkForwardOaS
 * with u outputs, srk_forward.cuh:468-501, then adjointSRK,
 * adjoint/srk_adjoint.cuh:402-425): the forward also writes u_x, u_y =
 * s conv(E, G) on the valid region [B x W x W] (FP32 in every precision), and
 * the adjoint takes them instead of recomputing u.  Costs 8 B W^2 bytes of
 * stored activations; saves the adjoint's convolution passes.
 */
int srkForwardBatchedU(const float* E, const float* I, const void* Gx_freq, const void* Gy_freq,
                       int N, int W, int erosion, float dx, int B, hipStream_t stream, float* R,
                       float* ux, float* uy, srk_workspace_t ws);
int srkAdjointBatchedU(const float* v, const float* ux, const float* uy, const float* I,
                       const void* Gx_freq, const void* Gy_freq, int N, int W, int erosion,
                       float dx, int B, hipStream_t stream, float* dLdI, float* dLdE,
                       srk_workspace_t ws);

/* Batched tangent / tangent-adjoint ([B x ...] arrays, as above). */
int srkTangentBatched(const float* f, const float* g, const float* df, const float* dg,
                      const void* Gx_freq, const void* Gy_freq, int N, int W, int erosion,
                      float dx, int B, hipStream_t stream, float* dh, srk_workspace_t ws);
int srkTangentAdjointBatched(const float* f, const float* g, const float* df, const float* dg,
                             const float* lam_h, const float* dlam_h, const void* Gx_freq,
                             const void* Gy_freq, int N, int W, int erosion, float dx, int B,
                             hipStream_t stream, float* dLam_f, float* dLam_g,
                             srk_workspace_t ws);

/* ---- reference test inputs (host) ---------------------------------------- */

/*
 * Replaces GpGenerator (third_party/include/stTestCore/gp_generator.h:95-201)
 * as used by TestPatternConfig::gaussian_process (third_party/src/patternUtil/
 * test_pattern.cu:156-160): count consecutive Gaussian-process fields of W x H
 * (column-major floats, written to the host array h_out [count][W*H]) with
 * covariance sigma^2 exp(-Q/2) of length scales ell_x, ell_y and correlation
 * rho, drawn from one std::mt19937_64 stream seeded with seed (the reference
 * replaces seed 0 by a random seed; here 0 is used as given).  Host only.
 */
int srkGpGenerate(double sigma, double ell_x, double ell_y, double rho, int W, int H,
                  unsigned long long seed, int count, float* h_out);

/* ---- critical-dimension (CD) measurement -------------------------------- */

/*
 * Replaces MeasureCD_Batched (third_party/src/metricUtil/cd_measure_util.cu:310-446,
 * declared at third_party/include/metricUtil/cd_measure_util.cuh:74-90): for B
 * row-major signals in [B x N] (device), threshold subtracted, upsample by L
 * with a 2A+1-tap sinc x Blackman polyphase filter, find the nearest crossings
 * left of and at-or-right of the centre (N L - 1)/2 and write, per signal,
 * left and right edge in pixels and cd = right - left (device arrays [B]; a
 * missing edge reads -FLT_MAX / L (left) or +FLT_MAX / L (right), finite, as
 * the reference's search start values, cd_measure_util.cu:37,218-219,264-282).
 * The reference returns a host vector and synchronises; this call is
 * asynchronous on the stream and allocates nothing.  threshold = 0 measures
 * zero crossings as the reference does.
 * Limits: N * L <= 2^24 and N + L(2A+1) + 512 floats within 160 KB of LDS.
 */
int srkMeasureCDBatched(const float* d_in, int B, int N, int L, int A, float threshold,
                        float* d_left_px, float* d_right_px, float* d_cd_px,
                        hipStream_t stream);

/*
 * Replaces BuildPolyphaseTable (cd_measure_util.cuh:11-18, cd_measure_util.cu:
 * 285-296): the sinc x Blackman polyphase table [L x (2A+1)], row-major, rows
 * normalised to sum 1, written to the caller's device buffer d_coeff (the
 * reference allocates it with cudaMallocAsync; this library never allocates
 * in a call).  Asynchronous.
 */
int srkBuildPolyphaseTable(int L, int A, float* d_coeff, hipStream_t stream);

/*
 * Replaces UpsampleBatched (cd_measure_util.cuh:20-33, cd_measure_util.cu:
 * 298-308): d_out[b][k L + r] = sum_i d_in[b][k + i - A] d_coeff[r][i], taps
 * outside the signal read zero; d_in [B x N], d_out [B x N L] row-major,
 * caller-allocated.  Asynchronous.
 */
int srkUpsampleBatched(const float* d_in, const float* d_coeff, int B, int N, int L, int A,
                       float* d_out, hipStream_t stream);

/*
 * Replaces the MeasureCD_Batched overload taking a prebuilt table and / or
 * upsampled signals (cd_measure_util.cuh:80-95, cd_measure_util.cu:310-332):
 * zero-crossing CD of B signals of length N.  d_coeff (nullable): the table
 * of srkBuildPolyphaseTable, else built internally; d_up (nullable): signals
 * already upsampled by srkUpsampleBatched [B x N L], else d_in [B x N] is
 * upsampled on the fly.  Outputs and missing edges as srkMeasureCDBatched.
 * Asynchronous, allocates nothing.
 */
int srkMeasureCDBatchedEx(const float* d_in, const float* d_coeff, const float* d_up, int B, int N,
                          int L, int A, float* d_left_px, float* d_right_px, float* d_cd_px,
                          hipStream_t stream);

/*
 * CD of the shrinkage result, fused: signal = E_crop + coeff_c * R on the centre
 * row (center_row != 0: row W/2 across the columns) or the centre column of
 * the W x W valid region, then as srkMeasureCDBatched with N = W.  E is
 * [B x N x N] and R [B x W x W], column-major; replaces crop_center +
 * apply_shrinkage + extract_center_row/col + MeasureCD_Batched of
 * training/main_cd_training.cu:128-165,213-261.
 */
int srkShrinkageCDBatched(const float* d_E, const float* d_R, int N, int W, int erosion, int B,
                          float coeff_c, int center_row, int L, int A, float threshold,
                          float* d_left_px, float* d_right_px, float* d_cd_px,
                          hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* SRK_AMD_H */
