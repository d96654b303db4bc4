// srk_cd.hip -- critical-dimension (CD) measurement on gfx950, one block per
// signal, for line/space checks of the shrinkage operator.
//
// Reference behaviour restated (file:line in the reference tree):
//   polyphase sinc x Blackman table, rows normalised to sum 1
//       third_party/src/metricUtil/cd_measure_util.cu:43-88
//   upsample out[k L + r] = sum_i sig[k + i - A] h[r][i], zero outside
//       cd_measure_util.cu:97-125
//   crossings between consecutive upsampled samples, x = i - y0 / (y1 - y0)
//       cd_measure_util.cu:134-170
//   nearest crossing left of / at-or-right of the centre (N L - 1) / 2, CD in
//       pixels = right / L - left / L        cd_measure_util.cu:205-282, 421
//   a missing edge keeps the search's start value -/+ FLT_MAX, so it reads
//       -/+ FLT_MAX / L pixels (finite)      cd_measure_util.cu:37,218-219,264-282
//   table build and batched upsampling as separate calls, and measurement from
//       a caller's table and / or upsampled signals
//                                           cd_measure_util.cu:285-308, 310-332
//   fused source E_crop + c R on the centre row / column
//       training/main_cd_training.cu:128-165
//
// The reference materialises the upsampled signals, a flag array, an exclusive
// scan and a compacted crossing list in device memory, then reduces.  Here a
// block keeps its signal in LDS, evaluates the upsampled samples on the fly
// over a contiguous run of edges per thread, and reduces the two nearest
// crossings in registers and LDS: no scratch memory, one launch for the batch.
#include "srk_cd.hpp"

#include <cmath>

namespace srk {

namespace {

constexpr int kCdThreads = 256;

// sinc(t) * Blackman window over [-A, A] (zero outside)
__device__ float sinc_blackman(float t, int A)
{
    if (fabsf(t) > (float)A + 1e-6f) return 0.f;
    const float s = fabsf(t) < 1e-6f ? 1.f : sinpif(t) / ((float)M_PI * t);
    const float u = (t + (float)A) / (float)(2 * A);
    return s * (0.42f - 0.5f * cospif(2.f * u) + 0.08f * cospif(4.f * u));
}

// Signal sources: value k of signal b, before the threshold is subtracted.
struct RowSource {
    const float* in;  // [B][N] row-major
    int N;
    __device__ float operator()(int b, int k) const { return in[(long long)b * N + k]; }
};

struct ShrinkSource {
    const float* E;  // [B][N][N] column-major
    const float* R;  // [B][W][W] column-major
    int N, W, e;
    float c;
    int center_row;
    __device__ float operator()(int b, int k) const
    {
        const int h = W / 2;
        const int r = center_row ? h : k, col = center_row ? k : h;
        const float ev = E[(long long)b * N * N + (long long)(col + e) * N + (r + e)];
        const float rv = R[(long long)b * W * W + (long long)col * W + r];
        return ev + c * rv;
    }
};

// row r (phase r / L) of the polyphase table, normalised to sum 1
__device__ void table_row(float* h, int r, int L, int A)
{
    const int taps = 2 * A + 1;
    const float phase = (float)r / (float)L;
    float sum = 0.f;
    for (int n = -A; n <= A; ++n) {
        const float wv = sinc_blackman((float)n - phase, A);
        h[n + A] = wv;
        sum += wv;
    }
    const float inv = fabsf(sum) > 1e-20f ? 1.f / sum : 1.f;
    for (int i = 0; i < taps; ++i) h[i] *= inv;
}

// start values of the nearest-crossing search; a missing edge keeps them
constexpr float kFltMax = 3.402823466e+38f;

// Block-wide reduction of the per-thread nearest crossings and the pixel
// conversion (x / L, then cd = right - left, all in FP32).
__device__ void finish_cd(float best_l, float best_r, float* red, int b, int L, float* left_px,
                          float* right_px, float* cd_px)
{
    red[threadIdx.x] = best_l;
    red[kCdThreads + threadIdx.x] = best_r;
    __syncthreads();
    for (int s = kCdThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
            red[kCdThreads + threadIdx.x] =
                fminf(red[kCdThreads + threadIdx.x], red[kCdThreads + threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float invL = 1.f / (float)L;
        const float lx = red[0] * invL, rx = red[kCdThreads] * invL;
        left_px[b] = lx;
        right_px[b] = rx;
        cd_px[b] = rx - lx;
    }
}

// crossing between samples i and i + 1 (values y0, y1): fold it into the
// nearest-left / nearest-right candidates
__device__ __forceinline__ void crossing(int i, float y0, float y1, float centre, float& best_l,
                                         float& best_r)
{
    if ((y0 <= 0.f && y1 > 0.f) || (y0 > 0.f && y1 <= 0.f)) {
        const float den = y1 - y0;
        const float x = fabsf(den) < 1e-30f ? (float)i : (float)i - y0 / den;
        if (x < centre)
            best_l = fmaxf(best_l, x);
        else
            best_r = fminf(best_r, x);
    }
}

// coeff: the caller's [L][2A+1] table, or null to build it in LDS
template <class Src>
__global__ __launch_bounds__(kCdThreads) void k_measure_cd(Src src, const float* coeff, int N,
                                                           int L, int A, float thr,
                                                           float* left_px, float* right_px,
                                                           float* cd_px)
{
    extern __shared__ float sm[];
    const int taps = 2 * A + 1;
    float* sig = sm;              // [N]
    float* h = sm + N;            // [L][taps]
    float* red = h + L * taps;    // [2][kCdThreads]
    const int b = blockIdx.x;
    for (int k = threadIdx.x; k < N; k += kCdThreads) sig[k] = src(b, k) - thr;
    if (coeff) {
        for (int i = threadIdx.x; i < L * taps; i += kCdThreads) h[i] = coeff[i];
    } else {
        for (int r = threadIdx.x; r < L; r += kCdThreads) table_row(h + r * taps, r, L, A);
    }
    __syncthreads();

    const int nup = N * L;
    const int nedge = nup - 1;
    const float centre = 0.5f * (float)(nup - 1);
    // upsampled sample m
    auto up = [&](int m) {
        const int k = m / L, r = m - k * L;
        const float* hr = h + r * taps;
        float acc = 0.f;
        for (int i = 0; i < taps; ++i) {
            const int idx = k + i - A;
            if ((unsigned)idx < (unsigned)N) acc += sig[idx] * hr[i];
        }
        return acc;
    };
    const int per = (nedge + kCdThreads - 1) / kCdThreads;
    const int i0 = threadIdx.x * per, i1 = min(nedge, i0 + per);
    float best_l = -kFltMax, best_r = kFltMax;
    if (i0 < i1) {
        float y0 = up(i0);
        for (int i = i0; i < i1; ++i) {
            const float y1 = up(i + 1);
            crossing(i, y0, y1, centre, best_l, best_r);
            y0 = y1;
        }
    }
    finish_cd(best_l, best_r, red, b, L, left_px, right_px, cd_px);
}

// CD from upsampled signals up [B][N L] (row-major): crossings of the given
// samples, no threshold (the reference measures zero crossings of them)
__global__ __launch_bounds__(kCdThreads) void k_measure_cd_up(const float* up, int N, int L,
                                                              float* left_px, float* right_px,
                                                              float* cd_px)
{
    __shared__ float red[2 * kCdThreads];
    const int b = blockIdx.x;
    const int nup = N * L, nedge = nup - 1;
    const float centre = 0.5f * (float)(nup - 1);
    const float* u = up + (long long)b * nup;
    const int per = (nedge + kCdThreads - 1) / kCdThreads;
    const int i0 = threadIdx.x * per, i1 = min(nedge, i0 + per);
    float best_l = -kFltMax, best_r = kFltMax;
    if (i0 < i1) {
        float y0 = u[i0];
        for (int i = i0; i < i1; ++i) {
            const float y1 = u[i + 1];
            crossing(i, y0, y1, centre, best_l, best_r);
            y0 = y1;
        }
    }
    finish_cd(best_l, best_r, red, b, L, left_px, right_px, cd_px);
}

// coeff[r][i], r < L: the polyphase table
__global__ void k_build_table(float* coeff, int L, int A)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < L) table_row(coeff + (long long)r * (2 * A + 1), r, L, A);
}

// out[b][k L + r] = sum_i in[b][k + i - A] coeff[r][i], zero outside the signal
__global__ void k_upsample(const float* in, const float* coeff, float* out, int N, int L, int A)
{
    const long long nup = (long long)N * L;
    const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (m >= nup) return;
    const int taps = 2 * A + 1;
    const int k = (int)(m / L), r = (int)(m - (long long)k * L);
    const float* sig = in + (long long)b * N;
    const float* h = coeff + (long long)r * taps;
    float acc = 0.f;
    for (int i = 0; i < taps; ++i) {
        const int idx = k + i - A;
        if ((unsigned)idx < (unsigned)N) acc += sig[idx] * h[i];
    }
    out[(long long)b * nup + m] = acc;
}

size_t cd_lds_bytes(int N, int L, int A)
{
    return ((size_t)N + (size_t)L * (2 * A + 1) + 2 * kCdThreads) * sizeof(float);
}

template <class Src>
int launch_cd(const Src& src, const float* coeff, int B, int N, int L, int A, float thr,
              float* l, float* r, float* cd, hipStream_t s)
{
    const size_t lds = cd_lds_bytes(N, L, A);
    // the dynamic-LDS limit is a per-kernel attribute: set once per process
    static hipError_t attr = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_measure_cd<Src>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)cd_max_lds_bytes());
    if (attr != hipSuccess) return (int)attr;
    k_measure_cd<Src><<<B, kCdThreads, lds, s>>>(src, coeff, N, L, A, thr, l, r, cd);
    return (int)hipGetLastError();
}

}  // namespace

size_t cd_max_lds_bytes() { return 160 * 1024; }

bool cd_shape_ok(int N, int L, int A)
{
    if (N < 2 || L < 1 || A < 1 || L > 4096 || A > 256) return false;
    if ((long long)N * L > (1LL << 24)) return false;  // sample indices exact in FP32
    return cd_lds_bytes(N, L, A) <= cd_max_lds_bytes();
}

int cd_measure_rows(const float* in, int B, int N, int L, int A, float thr, float* left_px,
                    float* right_px, float* cd_px, hipStream_t s)
{
    return launch_cd(RowSource{in, N}, nullptr, B, N, L, A, thr, left_px, right_px, cd_px, s);
}

int cd_measure_shrink(const float* E, const float* R, int N, int W, int e, int B, float c,
                      int center_row, int L, int A, float thr, float* left_px, float* right_px,
                      float* cd_px, hipStream_t s)
{
    return launch_cd(ShrinkSource{E, R, N, W, e, c, center_row}, nullptr, B, W, L, A, thr,
                     left_px, right_px, cd_px, s);
}

int cd_build_table(int L, int A, float* coeff, hipStream_t s)
{
    k_build_table<<<(L + 255) / 256, 256, 0, s>>>(coeff, L, A);
    return (int)hipGetLastError();
}

int cd_upsample(const float* in, const float* coeff, int B, int N, int L, int A, float* out,
                hipStream_t s)
{
    const long long nup = (long long)N * L;
    k_upsample<<<dim3((unsigned)((nup + 255) / 256), B), 256, 0, s>>>(in, coeff, out, N, L, A);
    return (int)hipGetLastError();
}

int cd_measure_ex(const float* in, const float* coeff, const float* up, int B, int N, int L,
                  int A, float* left_px, float* right_px, float* cd_px, hipStream_t s)
{
    if (up) {
        k_measure_cd_up<<<B, kCdThreads, 0, s>>>(up, N, L, left_px, right_px, cd_px);
        return (int)hipGetLastError();
    }
    return launch_cd(RowSource{in, N}, coeff, B, N, L, A, 0.f, left_px, right_px, cd_px, s);
}

}  // namespace srk
