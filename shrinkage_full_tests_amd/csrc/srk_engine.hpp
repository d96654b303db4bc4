// srk_engine.hpp -- host/device interface of the overlap-and-save tile engine.
//
// The engine evaluates the shrinkage operator family on M x M tiles in three
// passes, with exchange buffers between them sized for a chunk of tiles:
//
//   P1  (columns)  real FFT along y of every tile column (contiguous loads),
//                  written band-major: X1[tile][plane][k_y / RY][x][k_y % RY],
//                  so that each P2 block reads one contiguous band;
//   P2  (rows)     per k_y band: FFT along x, product with the kernel
//                  spectra, inverse FFT along x, crop to the needed columns:
//                  X2[tile][plane][x][k_y];
//   P3  (columns)  inverse FFT along y, fused with the gradient stencil, the
//                  combine / adjoint stencil and the valid-region crop; writes
//                  the caller's output directly.
//
// Spectra are held row-major by k_y (GT[d][k_y][k_x]) and pre-scaled by
// dx^2 / M^2, so P2 reads them coalesced along k_x.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace srk {

// Geometry of one tile, in global image coordinates (pointer free, so tile
// tables are built once per workspace).
struct Tile {
    int b;             // image index within the batch
    int p0x, p0y;      // window origin (column, row); may be negative
    int ox0, ox1;      // output columns [ox0, ox1)
    int oy0, oy1;      // output rows    [oy0, oy1)
    int pad;
};

// Arguments shared by every pass (passed by value).
struct EngineArgs {
    const float* g0;   // convolved field (g)           [B][N][N]
    const float* g1;   // second convolved field (dg)   [B][N][N] or null
    const float* f0;   // differentiated field (f)      [B][N][N]
    const float* f1;   // second differentiated field (df) or null
    const float* v0;   // cotangent (v, or d lambda_h)  [B][W][W]
    const float* v1;   // second cotangent (lambda_h)   [B][W][W] or null
    float* out0;       // h / dh [B][W][W], or lambda_f [B][N][N]
    float* out1;       // lambda_g [B][N][N]
    float* ux_out;     // optional u_x = s conv(g, Gx) on the valid region [B][W][W]
    float* uy_out;     // optional u_y
    long long img_stride;  // elements between images of N x N
    long long w_stride;    // elements between images of W x W
    int N, e, W;
    const float2* GT;      // [2][M/2+1][M] scaled spectra (Gx then Gy)
    const float2* tw;      // twiddles exp(-2 pi i k / M), k < M
    const float2* zcol;    // one all-zero column spectrum (spec_pitch(M) elements)
    const Tile* tiles;     // tiles of this chunk
    float2* X1;            // [chunk][np1] planes of M * HP elements, band layout (x1_index)
    float2* X2;            // [chunk][np2][S2][HP]
    int np1, np2;          // planes per tile in X1 / X2
    int S2;                // columns per plane in X2
    int halo2;             // extra stored columns on each side in X2 (0 fwd, 1 adj)
    int p2_rows;           // k_y bands per P2 block (tuning: 16, 8, 4; 0 = default)
    int ry_log2;           // log2 of the band height in effect (set by the engine)
    // P1A column blocks (4 tile columns each) [src_b0, src_b1) that hold
    // valid-region columns in some tile of the call; the P1A grid covers only
    // these (src_b1 <= src_b0: all M / 4)
    int src_b0, src_b1;
};

// Padded pitch (complex elements) of one column spectrum in X2, and the
// plane size M * spec_pitch(M) of both exchange buffers.
constexpr __host__ __device__ int spec_pitch(int M) { return M / 2 + 16; }

// X1 plane index of (k_y, x) in band layout with bands of 2^rs rows: bands of
// M * 2^rs elements, x-major inside a band.  The last band holds the Nyquist
// row k_y = M/2 alone, contiguous along x: M/2 * M + x.
template <int M>
__host__ __device__ constexpr int x1_index(int k, int x, int rs)
{
    return k == M / 2 ? M / 2 * M + x : (((k >> rs) * M + x) << rs) + (k & ((1 << rs) - 1));
}

enum class Op : int {
    Forward = 0,         // h = u(g) . grad f
    Tangent = 1,         // dh = u(dg) . grad f + u(g) . grad df
    Adjoint = 2,         // (lambda_f, lambda_g), u recomputed
    AdjointG = 3,        // lambda_g only (lambda_f from caller-provided u)
    TangentAdjoint = 4,  // (d lambda_f, d lambda_g)
};

// Kernels of the engine, for per-launch timing.
enum KernelKind : int {
    kP1Cols = 0,   // k_p1_cols
    kP1AdjSrc,     // k_p1_adj_src
    kP2Conv,       // k_p2_conv
    kP2Corr,       // k_p2_corr
    kP3Fwd,        // k_p3_fwd
    kP3Adj,        // k_p3_adj
    kFused,        // k_fused (the three passes of one operator in one persistent launch)
    kKernelKinds
};

// Host-side launch timer: when handed to engine_run, every launch is
// bracketed by two HIP events recorded on the launch stream.  Events are
// pooled and reused after collect().
struct LaunchTimer {
    struct Rec {
        int kind, tiles;
        hipEvent_t e0, e1;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    size_t used = 0;

    hipError_t record(hipStream_t s, hipEvent_t* ev);
    // Waits for the recorded launches and adds their durations (ms), counts
    // and tile counts into the per-kind arrays; then forgets them.
    hipError_t collect(double* ms, int* launches, long long* tiles);
    ~LaunchTimer();
};

// Second lane for chunk pipelining: chunks alternate between the caller's
// stream (exchange buffers args.X1 / args.X2) and `aux` (buffers X1b / X2b),
// so the passes of one chunk overlap those of the next.  `fork` / `join`
// order aux after the work already on the caller's stream and the caller's
// stream after aux (event record + stream wait, so it can be captured in a
// graph).
struct Fork {
    hipStream_t aux;
    hipEvent_t fork, join;
    float2* X1b;
    float2* X2b;
};

// Launch the passes of `op` for tiles [0, ntiles) of the tile table
// args.tiles, chunk tiles at a time.  prec 0: FP32 transforms; 1: FP16 MFMA
// DFT stages with FP32 accumulation (M <= 1024).  With `fork`, chunks
// alternate between two lanes.  Returns a hipError_t.
int engine_run(Op op, int M, int prec, const EngineArgs& args, int ntiles, int chunk,
               hipStream_t stream, LaunchTimer* timer = nullptr, const Fork* fork = nullptr);

// Fused pipeline (one persistent launch per operator): Op::Forward and
// Op::Adjoint in FP32 for one-tile images (N <= M, M = 256 or 512).  The
// exchange buffers args.X1 / args.X2 are rings of 8 * slots slots (np1 / np2
// planes each; slots per XCD in [lag + 1, kFusedSlotsMax], lag >= 1 rounds
// between stages); sync holds fused_sync_bytes() bytes of device memory.
// Returns a hipError_t.
constexpr int kFusedSlotsMax = 4;
bool fused_supported(Op op, int M, int N, int prec);
size_t fused_sync_bytes();
int engine_run_fused(Op op, int M, const EngineArgs& args, int B, void* sync, int slots, int lag,
                     hipStream_t stream, LaunchTimer* timer = nullptr);
// Timed-out waits of the last fused launch that used sync (waits for it).
int fused_sync_status(const void* sync, int* timeouts);

// Transpose + scale reference-layout spectra (Ghat[k_x*(M/2+1)+k_y]) into GT.
int engine_import_spectra(int M, const float2* Gx_ref, const float2* Gy_ref, float scale,
                          float2* GT, hipStream_t stream);

// Record a HIP error for srkLastHipError (the C ABI's per-thread last error).
void note_hip_error(int hip_error);

// lambda_f from caller-provided u (adjoint of the gradient stencil + combine).
int engine_adjoint_f_from_u(const float* v, const float* ux, const float* uy, float* lam_f,
                            int N, int W, int e, int B, long long img_stride,
                            long long w_stride, hipStream_t stream);


}  // namespace srk
