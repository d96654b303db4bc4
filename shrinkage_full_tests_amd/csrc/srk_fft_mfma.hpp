// srk_fft_mfma.hpp -- mixed-precision wavefront FFT on gfx950 matrix cores.
//
// Same contract as FftPlan<M> (srk_fft.hpp): one wavefront transforms one line
// of M complex values in LDS (padded slots, natural order in and out,
// unnormalised).  Plan: a radix-R0 Stockham pass on the vector ALUs straight
// from registers (R0 = M/256: 1, 2 or 4), then two radix-16 passes, each a
// dense DFT-matrix product on v_mfma_f32_16x16x32_f16:
//
//   rows = butterflies (16 per tile), K = (re, im) of the 16 inputs,
//   N = (re or im) of the 16 outputs (two products per tile),
//
// with FP16 operands, FP32 accumulation and FP32 twiddles.  SPLIT: every
// FP32 operand x and every DFT coefficient c is split into FP16 parts
// x = x_hi + x_lo, c = c_hi + c_lo, and each tile accumulates
// x_hi c_hi + x_lo c_hi + x_hi c_lo (three MFMAs): about 22 significant bits
// per operand instead of 11, so the transform error stays near FP32 level.
// Otherwise one product x_hi c_hi per tile (FP16 rounding of operands and
// coefficients, errors near 2^-11 of the line's magnitude).  Lane layout of
// the 16x16x32 instruction (checked on the device): A[i][k] in lane i + 16*(k/8),
// element k%8; B[k][n] in lane n + 16*(k/8), element k%8; D[i][n] in lane
// n + 16*(i/4), element i%4.  K index k = 8q + e: q = 0, 1 carry the real parts
// of inputs r = 8q + e, q = 2, 3 the imaginary parts of r = 8(q-2) + e.
//
// Twiddles of each pass are applied in FP32 to the outputs of the pass
// before it, so every MFMA operand is a plain real or imaginary part read
// from LDS.  Each line is scaled by a power of two first so that the FP16
// operands stay below 2^15 in magnitude; the scale is undone exactly on the
// final write.  The inverse transform is conj(DFT(conj(x))).
#pragma once

#include "srk_fft.hpp"

namespace srk {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4_t mfma16(const half8_t& a, const half8_t& b, const float4_t& c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// B operand tiles of the forward 16-point DFT, split into FP16 hi and lo
// parts: b[p][0] -> Re X_n, b[p][1] -> Im X_n (p = 0 hi, 1 lo; NP = 1 keeps
// the hi parts only).  The values come from a 64-lane table built on the host
// (srkWorkspaceCreate).
template <int NP>
struct Dft16Tiles {
    half8_t b[NP][2];

    __device__ void init(const half8_t* table, int lane)
    {
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) b[p][nb] = table[(lane * 2 + p) * 2 + nb];
    }
};

// One radix-16 MFMA Stockham pass with span Ns.  If !LAST the outputs are
// multiplied by tw_next (the next pass's twiddles, per lane); if LAST they are
// multiplied by `post` and conjugated when CONJ_OUT.
template <int M, int Ns, bool LAST, bool CONJ_OUT, bool SPLIT>
__device__ __forceinline__ void mfma16_pass(float2* __restrict__ line,
                                            const Dft16Tiles<SPLIT ? 2 : 1>& T,
                                            const float2* tw_next, float post, int lane)
{
    constexpr int NB = M / 16;    // butterflies
    constexpr int RB = NB / 16;   // row blocks
    const int i = lane & 15, q = lane >> 4;
    const bool im_part = q >= 2;
    half8_t Ahi[RB], Alo[SPLIT ? RB : 1];
    const float* lf = reinterpret_cast<const float*>(line) + (im_part ? 1 : 0);
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int j = 16 * rb + i;
            const int r = 8 * (q & 1) + e;
            const float v = lf[2 * lslot<M>(j + r * NB)];
            const _Float16 hi = (_Float16)v;
            Ahi[rb][e] = hi;
            if constexpr (SPLIT) Alo[rb][e] = (_Float16)(v - (float)hi);
        }
    float4_t Dre[RB], Dim[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        const float4_t z = {0.f, 0.f, 0.f, 0.f};
        if constexpr (SPLIT) {
            float4_t re = mfma16(Alo[rb], T.b[0][0], z);
            float4_t im = mfma16(Alo[rb], T.b[0][1], z);
            re = mfma16(Ahi[rb], T.b[1][0], re);
            im = mfma16(Ahi[rb], T.b[1][1], im);
            Dre[rb] = mfma16(Ahi[rb], T.b[0][0], re);
            Dim[rb] = mfma16(Ahi[rb], T.b[0][1], im);
        } else {
            Dre[rb] = mfma16(Ahi[rb], T.b[0][0], z);
            Dim[rb] = mfma16(Ahi[rb], T.b[0][1], z);
        }
    }
    // lane holds butterflies j = 16 rb + 4 q + e (e < 4), output s = i
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int j = 16 * rb + 4 * q + e;
            const int o = (j / Ns) * Ns * 16 + (j & (Ns - 1)) + i * Ns;
            float2 y = make_float2(Dre[rb][e], Dim[rb][e]);
            if constexpr (LAST) {
                y = make_float2(y.x * post, CONJ_OUT ? -y.y * post : y.y * post);
            } else {
                y = cmul(y, tw_next[rb * 4 + e]);
            }
            line[lslot<M>(o)] = y;
        }
}

template <int M, bool SPLIT = true>
struct FftPlanMfma {
    static constexpr int V = M / kWave;      // elements per lane
    static constexpr int R0 = M / 256;       // vector-ALU first pass radix (1 = none)
    static constexpr int NB16 = M / 16;      // butterflies of a radix-16 pass
    static constexpr int NOUT = NB16 / 16 * 4;  // outputs per lane of a radix-16 pass
    static constexpr int LOG_R0 = R0 == 1 ? 0 : R0 == 2 ? 1 : R0 == 4 ? 2 : 3;
    static_assert(M == 256 * R0 && (R0 == 1 || R0 == 2 || R0 == 4), "plan covers 256, 512, 1024");

    Dft16Tiles<SPLIT ? 2 : 1> T;
    float2 tw0[R0 > 1 ? V : 1];   // for outputs of the radix-R0 pass (next span R0)
    float2 twA[NOUT];             // for outputs of the first radix-16 pass (next span 16 R0)

    // tab: twiddles exp(-2 pi i k / M); dft16: per-lane DFT16 tiles (see Dft16Tiles),
    // stored right after the M twiddles.
    __device__ void load(const float2* tab, int lane)
    {
        T.init(reinterpret_cast<const half8_t*>(tab + M), lane);
        const int i = lane & 15, q = lane >> 4;
        if constexpr (R0 > 1) {
            constexpr int G = V / R0;
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int r = 0; r < R0; ++r) {
                    const int j = lane + kWave * g;
                    const int o = R0 * j + r;
                    const int jp = o % NB16, rp = o / NB16, kp = jp % R0;
                    tw0[g * R0 + r] = tab[(rp * kp * (M / (R0 * 16))) % M];
                }
        }
#pragma unroll
        for (int rb = 0; rb < NB16 / 16; ++rb)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int j = 16 * rb + 4 * q + e;
                const int o = (j / R0) * R0 * 16 + (j % R0) + i * R0;
                const int jp = o % NB16, rp = o / NB16, kp = jp % (16 * R0);
                twA[rb * 4 + e] = tab[(rp * kp * (M / (16 * R0 * 16))) % M];
            }
    }

    // Same contract as FftPlan<M>::run; with OREGS the result is read back from
    // LDS into out[] (the MFMA output layout is not the lane + 64 i layout).
    template <int DIR, bool REGS, bool OREGS>
    __device__ void run(float2* line, const float2 (&v)[V], float2 (&out)[V], int lane) const
    {
        run_lds<DIR, REGS>(line, v, lane);
        if constexpr (OREGS) {
#pragma unroll
            for (int k = 0; k < V; ++k) out[k] = line[lslot<M>(lane + kWave * k)];
        }
    }

    template <int DIR, bool REGS>
    __device__ void run_lds(float2* line, const float2 (&v)[V], int lane) const
    {
        float2 x[V];
#pragma unroll
        for (int k = 0; k < V; ++k) {
            x[k] = REGS ? v[k] : line[lslot<M>(lane + kWave * k)];
            if (DIR > 0) x[k].y = -x[k].y;   // inverse = conj(DFT(conj(.)))
        }
        // power-of-two scale: operands of the last MFMA pass stay below 2^15
        float m = 0.f;
#pragma unroll
        for (int k = 0; k < V; ++k) m = fmaxf(m, fmaxf(fabsf(x[k].x), fabsf(x[k].y)));
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
        int ex = 0;
        if (m > 0.f) {
            frexpf(m, &ex);
            ex -= 11 - LOG_R0;
        }
        const float sc = ldexpf(1.0f, -ex), post = ldexpf(1.0f, ex);
        if constexpr (R0 == 1) {
#pragma unroll
            for (int k = 0; k < V; ++k)
                line[lslot<M>(lane + kWave * k)] = make_float2(x[k].x * sc, x[k].y * sc);
        } else {
            // radix-R0 Stockham pass (span 1) on registers: butterfly j = lane + 64 g
            // takes elements j + r * (M/R0) = x[g + r*G]
            constexpr int G = V / R0;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                float2 y[R0];
#pragma unroll
                for (int r = 0; r < R0; ++r)
                    y[r] = make_float2(x[g + r * G].x * sc, x[g + r * G].y * sc);
                dft<R0, -1>(y);
                const int j = lane + kWave * g;
#pragma unroll
                for (int r = 0; r < R0; ++r)
                    line[lslot<M>(R0 * j + r)] = cmul(y[r], tw0[g * R0 + r]);
            }
        }
        mfma16_pass<M, R0, false, false, SPLIT>(line, T, twA, 1.0f, lane);
        if (DIR > 0)
            mfma16_pass<M, 16 * R0, true, true, SPLIT>(line, T, nullptr, post, lane);
        else
            mfma16_pass<M, 16 * R0, true, false, SPLIT>(line, T, nullptr, post, lane);
    }
};

}  // namespace srk
