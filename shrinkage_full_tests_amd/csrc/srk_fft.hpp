// srk_fft.hpp -- wavefront-level complex FFT for gfx950 (CDNA4).
//
// One 64-lane wavefront transforms one line of M complex values that lives in
// LDS.  The transform is a Stockham auto-sort FFT: every pass reads R values
// with stride M/R, applies twiddles, runs a radix-R DFT in registers and writes
// with stride Ns, so the result comes out in natural order without a
// bit-reversal step.  A pass first reads all of its operands, then writes, and
// only the owning wavefront touches the line, so passes run in place with no
// barrier (LDS operations of one wavefront complete in issue order).
//
// Twiddles come from a table tw[i] = exp(-2*pi*i*i/M) (i < M) built in double
// precision on the host; the inverse transform uses conj(tw).  Transforms are
// unnormalised in both directions.
//
// Radix plans (64 lanes): 256 = 4.4.4.4, 512 = 8.8.8, 1024 = 16.8.8,
// 2048 = 16.16.8.
#pragma once

#include <hip/hip_runtime.h>

namespace srk {

constexpr int kWave = 64;

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// cos / sin of 2*pi*j/16, j = 0..15 (exact constants of the 16th roots of unity)
__device__ constexpr float kCos16[16] = {
    1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
    0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
    -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
    0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f};
__device__ constexpr float kSin16[16] = {
    0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f,
    1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
    0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
    -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f};

// t * exp(DIR * 2*pi*i * J / L) with J, L compile-time; DIR = -1 forward.
template <int J, int L, int DIR>
__device__ __forceinline__ float2 twiddle_const(float2 t)
{
    if constexpr (J == 0) {
        return t;
    } else if constexpr (4 * J == L) {  // exp(DIR * i*pi/2) = DIR * i
        return DIR < 0 ? make_float2(t.y, -t.x) : make_float2(-t.y, t.x);
    } else {
        constexpr int idx = (J * 16 / L) & 15;
        const float c = kCos16[idx];
        const float s = DIR < 0 ? -kSin16[idx] : kSin16[idx];
        return make_float2(t.x * c - t.y * s, t.x * s + t.y * c);
    }
}

template <int R>
__device__ constexpr int bitrev(int i)
{
    int r = 0;
    for (int b = 1; b < R; b <<= 1) {
        r = (r << 1) | (i & 1);
        i >>= 1;
    }
    return r;
}

// Radix-2 stage of length L inside a radix-R register DFT (decimation in time).
template <int R, int L, int DIR, int J = 0>
__device__ __forceinline__ void dit_stage_j(float2 (&x)[R])
{
    if constexpr (J < L / 2) {
#pragma unroll
        for (int i = 0; i < R; i += L) {
            float2 t = twiddle_const<J, L, DIR>(x[i + J + L / 2]);
            float2 u = x[i + J];
            x[i + J] = cadd(u, t);
            x[i + J + L / 2] = csub(u, t);
        }
        dit_stage_j<R, L, DIR, J + 1>(x);
    }
}

template <int R, int L, int DIR>
__device__ __forceinline__ void dit_stages(float2 (&x)[R])
{
    if constexpr (L <= R) {
        dit_stage_j<R, L, DIR>(x);
        dit_stages<R, 2 * L, DIR>(x);
    }
}

// In-register DFT of size R (2, 4, 8, 16), natural order in and out.
template <int R, int DIR>
__device__ __forceinline__ void dft(float2 (&x)[R])
{
#pragma unroll
    for (int i = 0; i < R; ++i) {
        constexpr int dummy = 0;
        (void)dummy;
        const int j = bitrev<R>(i);
        if (i < j) {
            float2 t = x[i];
            x[i] = x[j];
            x[j] = t;
        }
    }
    dit_stages<R, 2, DIR>(x);
}

// One Stockham pass of radix R with span Ns over a line of M values.
// FROM_REGS: operands come from v[] (element lane + 64*i in v[i]) instead of LDS.
template <int M, int R, int Ns, int DIR, bool FROM_REGS>
__device__ __forceinline__ void stockham_pass(float2* __restrict__ line, const float2 (&v)[M / kWave],
                                              const float2* __restrict__ tw, int lane)
{
    constexpr int NB = M / R;       // butterflies in the pass
    constexpr int G = NB / kWave;   // butterflies per lane
    static_assert(NB % kWave == 0, "radix plan must give >= 64 butterflies per pass");
    float2 x[G][R];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if constexpr (FROM_REGS)
                x[g][r] = v[g + r * G];
            else
                x[g][r] = line[lane + kWave * g + r * NB];
        }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int j = lane + kWave * g;
        const int k = j & (Ns - 1);
        if constexpr (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 w = tw[r * k * (M / (Ns * R))];
                x[g][r] = DIR < 0 ? cmul(x[g][r], w) : cmulc(x[g][r], w);
            }
        }
        dft<R, DIR>(x[g]);
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) line[base + r * Ns] = x[g][r];
    }
}

template <int M>
struct FftPlan;
template <>
struct FftPlan<256> {
    template <int DIR, bool REGS>
    __device__ static void run(float2* line, const float2 (&v)[4], const float2* tw, int lane)
    {
        stockham_pass<256, 4, 1, DIR, REGS>(line, v, tw, lane);
        stockham_pass<256, 4, 4, DIR, false>(line, v, tw, lane);
        stockham_pass<256, 4, 16, DIR, false>(line, v, tw, lane);
        stockham_pass<256, 4, 64, DIR, false>(line, v, tw, lane);
    }
};
template <>
struct FftPlan<512> {
    template <int DIR, bool REGS>
    __device__ static void run(float2* line, const float2 (&v)[8], const float2* tw, int lane)
    {
        stockham_pass<512, 8, 1, DIR, REGS>(line, v, tw, lane);
        stockham_pass<512, 8, 8, DIR, false>(line, v, tw, lane);
        stockham_pass<512, 8, 64, DIR, false>(line, v, tw, lane);
    }
};
template <>
struct FftPlan<1024> {
    template <int DIR, bool REGS>
    __device__ static void run(float2* line, const float2 (&v)[16], const float2* tw, int lane)
    {
        stockham_pass<1024, 16, 1, DIR, REGS>(line, v, tw, lane);
        stockham_pass<1024, 8, 16, DIR, false>(line, v, tw, lane);
        stockham_pass<1024, 8, 128, DIR, false>(line, v, tw, lane);
    }
};
template <>
struct FftPlan<2048> {
    template <int DIR, bool REGS>
    __device__ static void run(float2* line, const float2 (&v)[32], const float2* tw, int lane)
    {
        stockham_pass<2048, 16, 1, DIR, REGS>(line, v, tw, lane);
        stockham_pass<2048, 16, 16, DIR, false>(line, v, tw, lane);
        stockham_pass<2048, 8, 256, DIR, false>(line, v, tw, lane);
    }
};

// Transform a line held in registers (element lane + 64*i in v[i]); result in
// `line` (LDS) in natural order.
template <int M, int DIR>
__device__ __forceinline__ void fft_from_regs(float2* line, const float2 (&v)[M / kWave],
                                              const float2* tw, int lane)
{
    FftPlan<M>::template run<DIR, true>(line, v, tw, lane);
}

// Transform a line already in LDS, in place.
template <int M, int DIR>
__device__ __forceinline__ void fft_in_lds(float2* line, const float2* tw, int lane)
{
    float2 dummy[M / kWave];
    FftPlan<M>::template run<DIR, false>(line, dummy, tw, lane);
}

}  // namespace srk
