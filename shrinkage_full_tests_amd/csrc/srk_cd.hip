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

template <class Src>
__global__ __launch_bounds__(kCdThreads) void k_measure_cd(Src src, int N, int L, int A,
                                                           float thr, float* left_px,
                                                           float* right_px, float* cd_px)
{
    extern __shared__ float sm[];
    const int taps = 2 * A + 1;
    float* sig = sm;              // [N]
    float* h = sm + N;            // [L][taps]
    float* red = h + L * taps;    // [2][kCdThreads]
    const int b = blockIdx.x;
    for (int k = threadIdx.x; k < N; k += kCdThreads) sig[k] = src(b, k) - thr;
    for (int r = threadIdx.x; r < L; r += kCdThreads) {
        const float phase = (float)r / (float)L;
        float sum = 0.f;
        for (int n = -A; n <= A; ++n) {
            const float wv = sinc_blackman((float)n - phase, A);
            h[r * taps + n + A] = wv;
            sum += wv;
        }
        const float inv = fabsf(sum) > 1e-20f ? 1.f / sum : 1.f;
        for (int i = 0; i < taps; ++i) h[r * taps + i] *= inv;
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
    float best_l = -INFINITY, best_r = INFINITY;
    if (i0 < i1) {
        float y0 = up(i0);
        for (int i = i0; i < i1; ++i) {
            const float y1 = up(i + 1);
            if ((y0 <= 0.f && y1 > 0.f) || (y0 > 0.f && y1 <= 0.f)) {
                const float den = y1 - y0;
                const float x = fabsf(den) < 1e-30f ? (float)i : (float)i - y0 / den;
                if (x < centre)
                    best_l = fmaxf(best_l, x);
                else
                    best_r = fminf(best_r, x);
            }
            y0 = y1;
        }
    }
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

size_t cd_lds_bytes(int N, int L, int A)
{
    return ((size_t)N + (size_t)L * (2 * A + 1) + 2 * kCdThreads) * sizeof(float);
}

template <class Src>
int launch_cd(const Src& src, int B, int N, int L, int A, float thr, float* l, float* r,
              float* cd, hipStream_t s)
{
    const size_t lds = cd_lds_bytes(N, L, A);
    static bool attr = false;
    if (!attr) {
        const hipError_t st = hipFuncSetAttribute(
            reinterpret_cast<const void*>(&k_measure_cd<Src>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)cd_max_lds_bytes());
        if (st != hipSuccess) return (int)st;
        attr = true;
    }
    k_measure_cd<Src><<<B, kCdThreads, lds, s>>>(src, N, L, A, thr, l, r, cd);
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
    return launch_cd(RowSource{in, N}, B, N, L, A, thr, left_px, right_px, cd_px, s);
}

int cd_measure_shrink(const float* E, const float* R, int N, int W, int e, int B, float c,
                      int center_row, int L, int A, float thr, float* left_px, float* right_px,
                      float* cd_px, hipStream_t s)
{
    return launch_cd(ShrinkSource{E, R, N, W, e, c, center_row}, B, W, L, A, thr, left_px,
                     right_px, cd_px, s);
}

}  // namespace srk
