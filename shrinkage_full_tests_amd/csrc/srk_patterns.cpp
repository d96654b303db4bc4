// srk_patterns.cpp -- host generators of the reference's test inputs.
//
// Gaussian-process fields restated from GpGenerator
// (third_party/include/stTestCore/gp_generator.h:95-201, used through
// TestPatternConfig::gaussian_process, third_party/src/patternUtil/
// test_pattern.cu:156-160): spectral synthesis on the power-of-two padded grid
//   c(dx, dy) = sigma^2 exp(-Q/2), Q = ((dx/lx)^2 - 2 rho (dx/lx)(dy/ly)
//                                        + (dy/ly)^2) / (1 - rho^2),
//   with nearest-image offsets on the padded grid,
//   S = FFT2(c), H = sqrt(max(0, Re S)),
//   per field: w ~ N(0, 1) (row-major over the padded grid), x = Re IFFT2(H FFT2 w),
//   output the top-left W x H block, column-major, as float.
// The noise stream is the reference's: std::mt19937_64 seeded with the seed and
// std::normal_distribution<double>, consumed field after field.  The FFT is
// this file's own (iterative radix-2 with an exact twiddle table), so values
// agree with the reference to double rounding before the final float cast.
#include "srk.h"

#include <cmath>
#include <complex>
#include <algorithm>
#include <random>
#include <thread>
#include <vector>

namespace {

using cd = std::complex<double>;

int pow2_at_least(int n)
{
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

// in-place FFT of length n (power of two); inverse divides by n
void fft_inplace(cd* a, int n, const std::vector<cd>& tw, bool inverse)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    const int N = (int)tw.size();  // twiddle table of the largest length
    for (int len = 2; len <= n; len <<= 1) {
        const int step = N / len;
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < len / 2; ++k) {
                cd w = tw[k * step];
                if (inverse) w = std::conj(w);
                const cd u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
    if (inverse)
        for (int i = 0; i < n; ++i) a[i] /= n;
}

// run fn(i) for i in [0, n) on up to 16 host threads (lines are independent,
// so the result does not depend on the thread count)
template <class Fn>
void parallel_lines(int n, Fn fn)
{
    const int T = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (T == 1 || n < 64) {
        for (int i = 0; i < n; ++i) fn(i);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&, t]() {
            for (int i = t; i < n; i += T) fn(i);
        });
    for (auto& x : th) x.join();
}

// 2-D FFT of a row-major w x h grid: rows, then columns
void fft2(std::vector<cd>& d, int w, int h, bool inverse)
{
    const int N = std::max(w, h);
    std::vector<cd> tw(N);
    for (int k = 0; k < N; ++k) tw[k] = std::polar(1.0, -2.0 * M_PI * k / N);
    parallel_lines(h, [&](int r) { fft_inplace(d.data() + (size_t)r * w, w, tw, inverse); });
    parallel_lines(w, [&](int c) {
        std::vector<cd> col(h);
        for (int r = 0; r < h; ++r) col[r] = d[(size_t)r * w + c];
        fft_inplace(col.data(), h, tw, inverse);
        for (int r = 0; r < h; ++r) d[(size_t)r * w + c] = col[r];
    });
}

}  // namespace

extern "C" int srkGpGenerate(double sigma, double ell_x, double ell_y, double rho, int W, int H,
                             unsigned long long seed, int count, float* h_out)
{
    if (!h_out) return SRK_ERR_NULL_PTR;
    if (!(sigma > 0) || !(ell_x > 0) || !(ell_y > 0) || !(rho > -1.0 && rho < 1.0))
        return SRK_ERR_INVALID_ARG;
    if (W < 1 || H < 1 || count < 1 || W > (1 << 15) || H > (1 << 15)) return SRK_ERR_BAD_DIMS;
    const int Wp = pow2_at_least(W), Hp = pow2_at_least(H);
    const size_t Np = (size_t)Wp * Hp;
    // amplitude filter from the wrapped covariance
    std::vector<cd> c(Np);
    for (int r = 0; r < Hp; ++r) {
        const double dy = r <= Hp / 2 ? r : r - Hp;
        for (int col = 0; col < Wp; ++col) {
            const double dx = col <= Wp / 2 ? col : col - Wp;
            const double u = dx / ell_x, v = dy / ell_y;
            const double q = (u * u - 2.0 * rho * u * v + v * v) / (1.0 - rho * rho);
            c[(size_t)r * Wp + col] = cd(sigma * sigma * std::exp(-0.5 * q), 0.0);
        }
    }
    fft2(c, Wp, Hp, false);
    std::vector<double> amp(Np);
    for (size_t i = 0; i < Np; ++i) amp[i] = std::sqrt(std::max(0.0, c[i].real()));

    std::mt19937_64 rng(static_cast<std::mt19937_64::result_type>(seed));
    std::normal_distribution<double> normal(0.0, 1.0);
    std::vector<cd> f(Np);
    for (int s = 0; s < count; ++s) {
        for (size_t i = 0; i < Np; ++i) f[i] = cd(normal(rng), 0.0);
        fft2(f, Wp, Hp, false);
        for (size_t i = 0; i < Np; ++i) f[i] *= amp[i];
        fft2(f, Wp, Hp, true);
        float* out = h_out + (size_t)s * W * H;
        for (int r = 0; r < H; ++r)
            for (int col = 0; col < W; ++col)
                out[(size_t)col * H + r] = (float)f[(size_t)r * Wp + col].real();
    }
    return SRK_OK;
}
