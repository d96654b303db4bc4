// srk_noalloc.hip -- the single-image forward (srkForwardOaS) for callers that
// cannot allocate device memory: code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc.
//
// Everything the workspace would hold (tile table, twiddles, the imported
// spectra, one all-zero column and the exchange buffers of a chunk of tiles)
// lives in `scratch`, a file-scope device array of kScratchBytes that every
// GPU the library runs on holds until the process ends.  Host code never
// passes the array's name to a kernel: its device address is resolved once
// per device with hipGetSymbolAddress (with that device current) and kept in
// g_scratch[device].  The call allocates nothing and synchronises nothing;
// tiles are streamed through the exchange buffers a chunk at a time, on the
// caller's stream only (no second lane).  Calls that overlap on one GPU share
// the array and race; each call rewrites every part of it that it reads.
#include "srk.h"

#include "srk_engine.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

int hip_fail(hipError_t e)
{
    srk::note_hip_error((int)e);
    return SRK_ERR_HIP;
}

// Device address of this device's scratch array (resolved on first use).
int scratch_for_current_device(void** out)
{
    int dev = 0;
    hipError_t st = hipGetDevice(&dev);
    if (st != hipSuccess) return hip_fail(st);
    if (dev < 0 || dev >= kMaxDevices) return SRK_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (st != hipSuccess) return hip_fail(st);
        if (!p) return hip_fail(hipErrorInvalidSymbol);
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return SRK_OK;
}

// Forward tile table of one image (the workspace's forward tiling): one
// window at the origin when N <= M, else windows of stride S = M - 2e
// (srk_forward.cuh:378-421), tile (tx, ty) at index ty * T + tx.
__global__ void k_fwd_tiles(srk::Tile* t, int N, int e, int M, int T)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T * T) return;
    srk::Tile x{};
    if (N <= M) {
        x = srk::Tile{0, 0, 0, e, N - e, e, N - e, 0};
    } else {
        const int S = M - 2 * e, tx = i % T, ty = i / T;
        const int ox0 = e + tx * S, oy0 = e + ty * S;
        x = srk::Tile{0, tx * S, ty * S, ox0, min(ox0 + S, N - e), oy0, min(oy0 + S, N - e), 0};
    }
    t[i] = x;
}

// twiddles tw[k] = exp(-2 pi i k / M) (FP64, rounded to FP32), then one
// all-zero column spectrum of `zc` elements
__global__ void k_tables(float2* tw, int M, float2* zcol, int zc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < M) {
        const double a = -2.0 * M_PI * k / M;
        tw[k] = make_float2((float)cos(a), (float)sin(a));
    }
    if (k < zc) zcol[k] = make_float2(0.f, 0.f);
}

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

extern "C" int srkForwardOaSNoAlloc(const float* E, const float* I, const void* Gx,
                                    const void* Gy, int N, int W, int erosion, float dx, int M,
                                    hipStream_t stream, float* R, float* ux, float* uy)
{
    if (!E || !I || !R || !Gx || !Gy) return SRK_ERR_NULL_PTR;
    if ((ux == nullptr) != (uy == nullptr)) return SRK_ERR_NULL_PTR;
    auto misaligned = [](const void* p, size_t a) { return (reinterpret_cast<size_t>(p) % a) != 0; };
    if (misaligned(E, 4) || misaligned(I, 4) || misaligned(R, 4) || misaligned(Gx, 8) ||
        misaligned(Gy, 8) || (ux && (misaligned(ux, 4) || misaligned(uy, 4))))
        return SRK_ERR_MISALIGNED;
    if (W != N - 2 * erosion || W <= 0 || erosion < 1) return SRK_ERR_BAD_DIMS;
    if ((M != 256 && M != 512 && M != 1024) || M <= 2 * erosion + 2)
        return SRK_ERR_UNSUPPORTED_SIZE;

    // geometry of the forward tiles
    const int T = N <= M ? 1 : (W + (M - 2 * erosion) - 1) / (M - 2 * erosion);
    const int ntiles = T * T;
    const int S2 = N <= M ? W : M - 2 * erosion;
    const int HP = srk::spec_pitch(M), H = M / 2 + 1;

    // scratch layout: tiles | twiddles | zero column | spectra (2 planes) | X1 | X2
    const size_t tiles_b = align_up((size_t)ntiles * sizeof(srk::Tile));
    const size_t tw_b = align_up((size_t)M * sizeof(float2));
    const size_t z_b = align_up((size_t)HP * sizeof(float2));
    const size_t gt_b = align_up((size_t)2 * H * M * sizeof(float2));
    const size_t x1_per = (size_t)M * HP * sizeof(float2);       // one plane (g)
    const size_t x2_per = (size_t)2 * S2 * HP * sizeof(float2);  // u_x, u_y
    const size_t fixed = tiles_b + tw_b + z_b + gt_b;
    if (fixed + x1_per + x2_per > kScratchBytes) return SRK_ERR_UNSUPPORTED_SIZE;
    const int chunk = (int)std::min<size_t>((size_t)ntiles, (kScratchBytes - fixed) / (x1_per + x2_per));

    void* base = nullptr;
    const int sc = scratch_for_current_device(&base);
    if (sc != SRK_OK) return sc;
    unsigned char* p = static_cast<unsigned char*>(base);
    srk::Tile* tiles = reinterpret_cast<srk::Tile*>(p);
    float2* tw = reinterpret_cast<float2*>(p + tiles_b);
    float2* zcol = reinterpret_cast<float2*>(p + tiles_b + tw_b);
    float2* GT = reinterpret_cast<float2*>(p + tiles_b + tw_b + z_b);
    float2* X1 = reinterpret_cast<float2*>(p + fixed);
    float2* X2 = reinterpret_cast<float2*>(p + fixed + (size_t)chunk * x1_per);

    k_fwd_tiles<<<(ntiles + 255) / 256, 256, 0, stream>>>(tiles, N, erosion, M, T);
    k_tables<<<(std::max(M, HP) + 255) / 256, 256, 0, stream>>>(tw, M, zcol, HP);
    hipError_t st = hipGetLastError();
    if (st != hipSuccess) return hip_fail(st);
    const float scale = (float)((double)dx * dx / ((double)M * M));
    st = (hipError_t)srk::engine_import_spectra(M, (const float2*)Gx, (const float2*)Gy, scale, GT,
                                                stream);
    if (st != hipSuccess) return hip_fail(st);

    srk::EngineArgs a{};
    a.g0 = E;
    a.f0 = I;
    a.out0 = R;
    a.ux_out = ux;
    a.uy_out = uy;
    a.N = N;
    a.e = erosion;
    a.W = W;
    a.img_stride = (long long)N * N;
    a.w_stride = (long long)W * W;
    a.GT = GT;
    a.tw = tw;
    a.zcol = zcol;
    a.tiles = tiles;
    a.X1 = X1;
    a.X2 = X2;
    a.np1 = 1;
    a.np2 = 2;
    a.S2 = S2;
    a.halo2 = 0;
    a.p2_rows = 0;
    st = (hipError_t)srk::engine_run(srk::Op::Forward, M, 0, a, ntiles, chunk, stream);
    return st == hipSuccess ? SRK_OK : hip_fail(st);
}
