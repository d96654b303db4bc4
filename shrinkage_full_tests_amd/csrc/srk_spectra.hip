// srk_spectra.hip -- SRSRK kernel spectra generated on the device in FP64.
//
// Same mathematics as the host path of srkKernelSpectraSRSRK (srk_api.cpp):
// the Green's tensor Gx, Gy = coeff K1(gamma r) Blackman(r / Rw) (x/r, y/r)
// of training/main_cd_training.cu:63-77 and tan_tanadj/main_tan_tanadj.cu:93-116,
// K1 from the ascending series (A&S 9.6.11) for x <= 2 and the integral
// representation (A&S 9.6.24) beyond, and the 2-D DFT of the centred kernel at
// tile size M in the reference spectrum layout out[k_x (M/2+1) + k_y]
// (srk_forward.cuh:154-203: zero pad, fftshift to the origin, R2C).  Every
// step runs in FP64 on the device; only the final spectra are rounded to
// FP32.  The DFT is separable: first along y into T[d][i][k_y], then along x.
#include "srk_spectra.hpp"

#include <algorithm>
#include <cmath>

namespace srk {

namespace {

__device__ double k1_series(double x)
{
    const double euler = 0.57721566490153286061;
    const double y = 0.25 * x * x;
    double term = 1.0, psi1 = -euler, psi2 = 1.0 - euler, si = 0.0, s2 = 0.0;
    for (int k = 0; k < 80; ++k) {
        si += term;
        s2 += (psi1 + psi2) * term;
        term *= y / ((k + 1.0) * (k + 2.0));
        psi1 += 1.0 / (k + 1.0);
        psi2 += 1.0 / (k + 2.0);
        if (term < 1e-20 * si) break;
    }
    const double i1 = 0.5 * x * si;
    return 1.0 / x + log(0.5 * x) * i1 - 0.25 * x * s2;
}

__device__ double k1_integral(double x)
{
    const double h = 0.0625;
    double acc = 0.5 * exp(-x);
    for (int k = 1; k < 20000; ++k) {
        const double c = cosh(k * h);
        const double v = exp(-x * c) * c;
        acc += v;
        if (v < 1e-20 * acc) break;
    }
    return h * acc;
}

__device__ double bessel_k1(double x) { return x <= 2.0 ? k1_series(x) : k1_integral(x); }

// K[d][i * kw + j] (i = x / column index, j = y / row index), FP64 and FP32
__global__ void k_srsrk_spatial(int kw, double sigma, double dx, double cw, double* K64,
                                float* kx32, float* ky32)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= kw * kw) return;
    const int i = idx / kw, j = idx - i * kw, c = kw / 2;
    const double gamma = 1.0 / sigma, Rw = cw * sigma, coeff = gamma / (2.0 * M_PI);
    const double x = (i - c) * dx, y = (j - c) * dx;
    const double r = sqrt(x * x + y * y);
    double gxv = 0.0, gyv = 0.0;
    if (!(r < 1e-6 * dx || r > Rw)) {
        const double t = r / Rw;
        const double win = 0.42 + 0.5 * cospi(t) + 0.08 * cospi(2.0 * t);
        const double base = coeff * bessel_k1(gamma * r) * win;
        gxv = base * (x / r);
        gyv = base * (y / r);
    }
    K64[idx] = gxv;
    K64[kw * kw + idx] = gyv;
    if (kx32) kx32[idx] = (float)gxv;
    if (ky32) ky32[idx] = (float)gyv;
}

// tw[n] = exp(-2 pi i n / M)
__global__ void k_twiddles64(int M, double2* tw)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= M) return;
    double s, c;
    sincospi(-2.0 * n / M, &s, &c);
    tw[n] = make_double2(c, s);
}

__device__ __forceinline__ int wrap(long long v, int M)
{
    long long r = v % M;
    return (int)(r < 0 ? r + M : r);
}

// T[d][i][ky] = sum_j K[d][i][j] exp(-2 pi i ky (j - c) / M)
__global__ void k_dft_y(const double* K64, int kw, int M, const double2* tw, double2* T)
{
    const int H = M / 2 + 1;
    const int ky = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y, d = blockIdx.z;
    if (ky >= H) return;
    const double* g = K64 + (size_t)d * kw * kw + (size_t)i * kw;
    const int c = kw / 2;
    double re = 0.0, im = 0.0;
    for (int j = 0; j < kw; ++j) {
        const double2 w = tw[wrap((long long)ky * (j - c), M)];
        re += g[j] * w.x;
        im += g[j] * w.y;
    }
    T[((size_t)d * kw + i) * H + ky] = make_double2(re, im);
}

// out_d[kx H + ky] = sum_i exp(-2 pi i kx (i - c) / M) T[d][i][ky], rounded to FP32
__global__ void k_dft_x(const double2* T, int kw, int M, const double2* tw, float2* gx,
                        float2* gy)
{
    const int H = M / 2 + 1;
    const int ky = blockIdx.x * blockDim.x + threadIdx.x;
    const int kx = blockIdx.y, d = blockIdx.z;
    if (ky >= H) return;
    const double2* t = T + (size_t)d * kw * H + ky;
    const int c = kw / 2;
    double re = 0.0, im = 0.0;
    for (int i = 0; i < kw; ++i) {
        const double2 w = tw[wrap((long long)kx * (i - c), M)];
        const double2 v = t[(size_t)i * H];
        re += w.x * v.x - w.y * v.y;
        im += w.x * v.y + w.y * v.x;
    }
    (d == 0 ? gx : gy)[(size_t)kx * H + ky] = make_float2((float)re, (float)im);
}

// ---------------------------------------------------------------------------
// Spectra handed over at another size.  A reference-layout half spectrum of
// size Ns (Ghat[k_x (Ns/2+1) + k_y]) fixes the Ns-periodic spatial kernel
//   K(x, y) = Ns^-2 sum_{k_x, k_y} Ghat(k_x, k_y) exp(+2 pi i (k_x x + k_y y) / Ns);
// only its values on the erosion box |x|, |y| <= e enter the valid region, so
// they are evaluated directly (x first, then y with the Hermitian weights:
// k_y = 0 and the Nyquist row once, the other rows twice, real part).
// ---------------------------------------------------------------------------

// T[d][i][k_y] = sum_kx G_d[kx H + ky] exp(+2 pi i kx (i - e) / Ns)
__global__ void k_inv_x(const float2* __restrict__ g0, const float2* __restrict__ g1, int Ns,
                        int e, const double2* __restrict__ tw, double2* __restrict__ T)
{
    const int H = Ns / 2 + 1, kw = 2 * e + 1;
    const int ky = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y, d = blockIdx.z;
    if (ky >= H) return;
    const float2* g = (d == 0 ? g0 : g1) + ky;
    const long long x = i - e;
    double re = 0.0, im = 0.0;
    for (int kx = 0; kx < Ns; ++kx) {
        const double2 w = tw[wrap(kx * x, Ns)];  // exp(-2 pi i kx x / Ns); conjugated below
        const float2 v = g[(size_t)kx * H];
        re += v.x * w.x + v.y * w.y;
        im += v.y * w.x - v.x * w.y;
    }
    T[((size_t)d * kw + i) * H + ky] = make_double2(re, im);
}

// K64[d][i kw + j] = Ns^-2 sum_ky c(ky) Re(T[d][i][ky] exp(+2 pi i ky (j - e) / Ns))
__global__ void k_inv_y(const double2* __restrict__ T, int Ns, int e,
                        const double2* __restrict__ tw, double* __restrict__ K64)
{
    const int H = Ns / 2 + 1, kw = 2 * e + 1;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y, d = blockIdx.z;
    if (j >= kw) return;
    const double2* t = T + ((size_t)d * kw + i) * H;
    const long long y = j - e;
    double acc = 0.0;
    for (int ky = 0; ky < H; ++ky) {
        const double2 w = tw[wrap(ky * y, Ns)];
        const double c = (ky == 0 || 2 * ky == Ns) ? 1.0 : 2.0;
        acc += c * (t[ky].x * w.x + t[ky].y * w.y);  // Re(t conj(w))
    }
    K64[(size_t)d * kw * kw + (size_t)i * kw + j] = acc / ((double)Ns * Ns);
}

// Partial energies: part[2 b] = sum over this block's spectrum bins of
// c(ky) |G|^2 / Ns^2 (Parseval: the whole kernel's energy), part[2 b + 1] =
// sum of the box values' squares.
__global__ void k_energy_part(const float2* __restrict__ g0, const float2* __restrict__ g1, int Ns,
                              const double* __restrict__ K64, int kw, double* __restrict__ part)
{
    __shared__ double s0[256], s1[256];
    const int H = Ns / 2 + 1;
    const long long nspec = 2LL * Ns * H, nbox = 2LL * kw * kw;
    double a = 0.0, b = 0.0;
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < nspec; q += 256LL * gridDim.x) {
        const long long d = q / ((long long)Ns * H), r = q - d * Ns * H;
        const int ky = (int)(r % H);
        const float2 v = (d == 0 ? g0 : g1)[r];
        const double c = (ky == 0 || 2 * ky == Ns) ? 1.0 : 2.0;
        a += c * ((double)v.x * v.x + (double)v.y * v.y);
    }
    for (long long q = blockIdx.x * 256LL + threadIdx.x; q < nbox; q += 256LL * gridDim.x)
        b += K64[q] * K64[q];
    s0[threadIdx.x] = a / ((double)Ns * Ns);
    s1[threadIdx.x] = b;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s0[threadIdx.x] += s0[threadIdx.x + h];
            s1[threadIdx.x] += s1[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0[0];
        part[2 * blockIdx.x + 1] = s1[0];
    }
}

// ratio = sqrt(max(0, E_all - E_box) / E_all): relative RMS of the kernel
// outside the erosion box (0 for an all-zero spectrum)
__global__ void k_energy_final(const double* __restrict__ part, int nb, double* __restrict__ ratio)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0, b = 0.0;
    for (int i = 0; i < nb; ++i) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    *ratio = a > 0.0 ? sqrt(fmax(0.0, a - b) / a) : 0.0;
}

constexpr int kEnergyBlocks = 256;

}  // namespace

size_t spectra_scratch_bytes(int kw, int M)
{
    const size_t H = M / 2 + 1;
    return 2 * (size_t)kw * kw * sizeof(double) + (size_t)M * sizeof(double2) +
           2 * (size_t)kw * H * sizeof(double2);
}

int spectra_srsrk_device(int kw, double sigma, double dx, double cw, int M, void* scratch,
                         float2* gx, float2* gy, float* kx32, float* ky32, hipStream_t s)
{
    const int H = M / 2 + 1;
    double* K64 = static_cast<double*>(scratch);
    double2* tw = reinterpret_cast<double2*>(K64 + 2 * (size_t)kw * kw);
    double2* T = tw + M;
    k_srsrk_spatial<<<(kw * kw + 255) / 256, 256, 0, s>>>(kw, sigma, dx, cw, K64, kx32, ky32);
    k_twiddles64<<<(M + 255) / 256, 256, 0, s>>>(M, tw);
    k_dft_y<<<dim3((H + 127) / 128, kw, 2), 128, 0, s>>>(K64, kw, M, tw, T);
    k_dft_x<<<dim3((H + 127) / 128, M, 2), 128, 0, s>>>(T, kw, M, tw, gx, gy);
    return (int)hipGetLastError();
}

}  // namespace srk

namespace srk {

size_t resample_scratch_bytes(int Ns_max, int e, int M)
{
    const size_t kw = 2 * (size_t)e + 1, Hs = Ns_max / 2 + 1, Hm = M / 2 + 1;
    return 2 * kw * kw * sizeof(double) +                    // K64
           (size_t)std::max(Ns_max, M) * sizeof(double2) +   // twiddles (one size at a time)
           2 * kw * std::max(Hs, Hm) * sizeof(double2) +     // T (inverse, then forward DFT)
           2 * (size_t)kEnergyBlocks * sizeof(double) +      // energy partials
           sizeof(double);                                   // ratio
}

int resample_spectra_device(int Ns, int e, int M, const float2* gx_src, const float2* gy_src,
                            void* scratch, float2* gx_dst, float2* gy_dst, double** ratio,
                            int Ns_max, hipStream_t s)
{
    const int kw = 2 * e + 1, Hs = Ns / 2 + 1, Hm = M / 2 + 1;
    double* K64 = static_cast<double*>(scratch);
    double2* tw = reinterpret_cast<double2*>(K64 + 2 * (size_t)kw * kw);
    double2* T = tw + std::max(Ns_max, M);
    double* part = reinterpret_cast<double*>(
        T + 2 * (size_t)kw * std::max<size_t>(Ns_max / 2 + 1, Hm));
    double* r = part + 2 * kEnergyBlocks;
    *ratio = r;
    k_twiddles64<<<(Ns + 255) / 256, 256, 0, s>>>(Ns, tw);
    k_inv_x<<<dim3((Hs + 127) / 128, kw, 2), 128, 0, s>>>(gx_src, gy_src, Ns, e, tw, T);
    k_inv_y<<<dim3((kw + 63) / 64, kw, 2), 64, 0, s>>>(T, Ns, e, tw, K64);
    k_energy_part<<<kEnergyBlocks, 256, 0, s>>>(gx_src, gy_src, Ns, K64, kw, part);
    k_energy_final<<<1, 64, 0, s>>>(part, kEnergyBlocks, r);
    if (gx_dst && gy_dst) {
        k_twiddles64<<<(M + 255) / 256, 256, 0, s>>>(M, tw);
        k_dft_y<<<dim3((Hm + 127) / 128, kw, 2), 128, 0, s>>>(K64, kw, M, tw, T);
        k_dft_x<<<dim3((Hm + 127) / 128, M, 2), 128, 0, s>>>(T, kw, M, tw, gx_dst, gy_dst);
    }
    return (int)hipGetLastError();
}

}  // namespace srk
