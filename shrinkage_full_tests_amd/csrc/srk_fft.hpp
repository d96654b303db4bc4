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
// LDS line layout: element i lives at slot i + i/8 (one padding slot per
// eight).  Without it the first two radix-8 passes write with a stride of 8 or
// 64 complex values and every 64-bit store of a wavefront lands on the same
// few banks.
//
// Twiddles: a lane's twiddles depend only on its butterfly index, which is the
// same for every line, so each kernel loads them once (FftPlan::load) from a
// table tw[i] = exp(-2*pi*i*i/M) built in double precision on the host and
// keeps them in registers; the inverse uses their conjugates.  Transforms are
// unnormalised in both directions.
//
// Radix plans (64 lanes): 256 = 4.4.4.4, 512 = 8.8.8, 1024 = 16.8.8,
// 2048 = 16.16.8.
#pragma once

#include <hip/hip_runtime.h>

namespace srk {

constexpr int kWave = 64;

// Padded layout of an LDS line of M elements: one pad slot after every 2^S
// elements, S = pad_shift<M>().  Measured LDS banking (reads: 64 four-byte
// banks, 32 lanes a cycle; writes: 32 banks, 16 lanes a cycle) favours a pad
// every 8 elements for the 512-point plan (8^3) and every 16 for the
// 1024- and 2048-point plans, whose radix-16 first pass writes at stride 16.
template <int M>
__host__ __device__ constexpr int pad_shift() { return M >= 1024 ? 4 : 3; }
// slot of element i in an LDS line, and slots per line of M elements
template <int M>
__host__ __device__ constexpr int lslot(int i) { return i + (i >> pad_shift<M>()); }
__host__ __device__ constexpr int line_slots(int M) { return M >= 1024 ? M + M / 16 : M + M / 8; }

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
        const int j = bitrev<R>(i);
        if (i < j) {
            float2 t = x[i];
            x[i] = x[j];
            x[j] = t;
        }
    }
    dit_stages<R, 2, DIR>(x);
}

// Per-lane twiddles of one pass.  For Ns <= 64 the twiddle index k = j mod Ns
// of butterfly j = lane + 64 g does not depend on g, so one set is kept;
// otherwise one set per g.  CACHE = false reads the table at use instead.
template <int M, int R, int Ns, bool CACHE = true>
struct PassTw {
    static constexpr int NB = M / R;
    static constexpr int G = NB / kWave;
    static constexpr bool kShared = Ns <= kWave;
    static constexpr int NG = (Ns == 1 || !CACHE) ? 0 : (kShared ? 1 : G);
    float2 w[NG > 0 ? NG : 1][R - 1];
    const float2* tab;

    __device__ __forceinline__ void load(const float2* tw, int lane)
    {
        tab = tw;
        if constexpr (NG > 0) {
#pragma unroll
            for (int gi = 0; gi < NG; ++gi) {
                const int k = (lane + kWave * gi) & (Ns - 1);
#pragma unroll
                for (int r = 1; r < R; ++r) w[gi][r - 1] = tw[r * k * (M / (Ns * R))];
            }
        }
    }
    __device__ __forceinline__ float2 get(int g, int r, int lane) const
    {
        if constexpr (NG > 0) {
            return w[kShared ? 0 : g][r - 1];
        } else {
            const int k = (lane + kWave * g) & (Ns - 1);
            return tab[r * k * (M / (Ns * R))];
        }
    }
};

// Slot arithmetic of the padded layout.  For C a multiple of the pad period
// P = 2^S, lslot(a + C) = lslot(a) + C + C/P, so every access of a pass is a
// per-lane base slot plus a compile-time offset (immediate LDS offsets, no
// per-access address arithmetic).
template <int M>
__host__ __device__ constexpr int slot_off(int C) { return C + (C >> pad_shift<M>()); }

// Base slot and constant offset of output r of butterfly j in a pass of radix R
// and span Ns: element e = A + B + r Ns with A = (j/Ns) Ns R, B = j mod Ns.
// For Ns a multiple of P, lslot(e) = lslot(A + B) + slot_off(r Ns).  For
// Ns < P (Ns divides P), A + B + r Ns never crosses a pad boundary beyond the
// ones r Ns crosses by itself: lslot(e) = A + A/P + B + slot_off(r Ns).
template <int M, int R, int Ns>
struct WriteSlots {
    static constexpr int S = pad_shift<M>();
    __device__ static __forceinline__ int base(int j)
    {
        const int A = (j / Ns) * Ns * R, B = j & (Ns - 1);
        if constexpr (Ns % (1 << S) == 0)
            return lslot<M>(A + B);
        else
            return A + (A >> S) + B;
    }
    __device__ static __forceinline__ int offset(int r) { return slot_off<M>(r * Ns); }
};

// One Stockham pass of radix R with span Ns over a line of M values.
// FROM_REGS: operands come from v[] (element lane + 64*i in v[i]) instead of
// LDS.  TO_REGS (last pass only, Ns * R = M): outputs stay in registers,
// element lane + 64*i in out[i], instead of being written to LDS.
template <int M, int R, int Ns, int DIR, bool FROM_REGS, bool TO_REGS, bool CACHE>
__device__ __forceinline__ void stockham_pass(float2* __restrict__ line, const float2 (&v)[M / kWave],
                                              float2 (&out)[M / kWave],
                                              const PassTw<M, R, Ns, CACHE>& T, int lane)
{
    constexpr int NB = M / R;       // butterflies in the pass
    constexpr int G = NB / kWave;   // butterflies per lane
    static_assert(NB % kWave == 0, "radix plan must give >= 64 butterflies per pass");
    static_assert(!TO_REGS || Ns * R == M, "only the last pass can leave its outputs in registers");
    float2 x[G][R];
    if constexpr (FROM_REGS) {
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int r = 0; r < R; ++r) x[g][r] = v[g + r * G];
    } else {
        const float2* rb = line + lslot<M>(lane);
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int r = 0; r < R; ++r) x[g][r] = rb[slot_off<M>(kWave * g + r * NB)];
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int j = lane + kWave * g;
        if constexpr (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const float2 w = T.get(g, r, lane);
                x[g][r] = DIR < 0 ? cmul(x[g][r], w) : cmulc(x[g][r], w);
            }
        }
        dft<R, DIR>(x[g]);
        if constexpr (TO_REGS) {
            // output r of butterfly j is element j + r * NB = lane + 64 (g + r G)
#pragma unroll
            for (int r = 0; r < R; ++r) out[g + r * G] = x[g][r];
        } else {
            float2* wb = line + WriteSlots<M, R, Ns>::base(j);
#pragma unroll
            for (int r = 0; r < R; ++r) wb[WriteSlots<M, R, Ns>::offset(r)] = x[g][r];
        }
    }
}

template <int M>
struct FftPlan;

// run<DIR, FROM_REGS, TO_REGS>(line, v, out, lane): the transform of the line
// (from v[] if FROM_REGS, else from LDS), result in LDS in natural order, or in
// out[] (element lane + 64*i in out[i]) if TO_REGS.  Unnormalised.
template <>
struct FftPlan<256> {
    static constexpr int V = 256 / kWave;
    PassTw<256, 4, 1> t0;
    PassTw<256, 4, 4> t1;
    PassTw<256, 4, 16> t2;
    PassTw<256, 4, 64> t3;
    __device__ void load(const float2* tw, int lane)
    {
        t0.load(tw, lane); t1.load(tw, lane); t2.load(tw, lane); t3.load(tw, lane);
    }
    template <int DIR, bool REGS, bool OREGS>
    __device__ void run(float2* line, const float2 (&v)[V], float2 (&o)[V], int lane) const
    {
        stockham_pass<256, 4, 1, DIR, REGS, false>(line, v, o, t0, lane);
        stockham_pass<256, 4, 4, DIR, false, false>(line, v, o, t1, lane);
        stockham_pass<256, 4, 16, DIR, false, false>(line, v, o, t2, lane);
        stockham_pass<256, 4, 64, DIR, false, OREGS>(line, v, o, t3, lane);
    }
};
template <>
struct FftPlan<512> {
    static constexpr int V = 512 / kWave;
    PassTw<512, 8, 1> t0;
    PassTw<512, 8, 8> t1;
    PassTw<512, 8, 64> t2;
    __device__ void load(const float2* tw, int lane)
    {
        t0.load(tw, lane); t1.load(tw, lane); t2.load(tw, lane);
    }
    template <int DIR, bool REGS, bool OREGS>
    __device__ void run(float2* line, const float2 (&v)[V], float2 (&o)[V], int lane) const
    {
        stockham_pass<512, 8, 1, DIR, REGS, false>(line, v, o, t0, lane);
        stockham_pass<512, 8, 8, DIR, false, false>(line, v, o, t1, lane);
        stockham_pass<512, 8, 64, DIR, false, OREGS>(line, v, o, t2, lane);
    }
};
template <>
struct FftPlan<1024> {
    static constexpr int V = 1024 / kWave;
    PassTw<1024, 16, 1> t0;
    PassTw<1024, 8, 16> t1;
    PassTw<1024, 8, 128> t2;
    __device__ void load(const float2* tw, int lane)
    {
        t0.load(tw, lane); t1.load(tw, lane); t2.load(tw, lane);
    }
    template <int DIR, bool REGS, bool OREGS>
    __device__ void run(float2* line, const float2 (&v)[V], float2 (&o)[V], int lane) const
    {
        stockham_pass<1024, 16, 1, DIR, REGS, false>(line, v, o, t0, lane);
        stockham_pass<1024, 8, 16, DIR, false, false>(line, v, o, t1, lane);
        stockham_pass<1024, 8, 128, DIR, false, OREGS>(line, v, o, t2, lane);
    }
};
template <>
struct FftPlan<2048> {
    static constexpr int V = 2048 / kWave;
    PassTw<2048, 16, 1> t0;
    PassTw<2048, 16, 16> t1;
    PassTw<2048, 8, 256, false> t2;
    __device__ void load(const float2* tw, int lane)
    {
        t0.load(tw, lane); t1.load(tw, lane); t2.load(tw, lane);
    }
    template <int DIR, bool REGS, bool OREGS>
    __device__ void run(float2* line, const float2 (&v)[V], float2 (&o)[V], int lane) const
    {
        stockham_pass<2048, 16, 1, DIR, REGS, false>(line, v, o, t0, lane);
        stockham_pass<2048, 16, 16, DIR, false, false>(line, v, o, t1, lane);
        stockham_pass<2048, 8, 256, DIR, false, OREGS>(line, v, o, t2, lane);
    }
};

// Helpers over any plan PL with the interface above (FftPlan<M>, FftPlanMfma<M>).
// Register arrays hold element lane + 64*i in index i; `line` is the LDS line
// of the calling wavefront (input for the _lds forms, scratch otherwise).

// registers -> LDS (natural order)
template <int M, int DIR, class PL>
__device__ __forceinline__ void fft_from_regs(const PL& P, float2* line,
                                              const float2 (&v)[M / kWave], int lane)
{
    float2 unused[M / kWave];
    P.template run<DIR, true, false>(line, v, unused, lane);
}

// LDS -> LDS, in place
template <int M, int DIR, class PL>
__device__ __forceinline__ void fft_in_lds(const PL& P, float2* line, int lane)
{
    float2 unused[M / kWave];
    P.template run<DIR, false, false>(line, unused, unused, lane);
}

// LDS -> registers
template <int M, int DIR, class PL>
__device__ __forceinline__ void fft_lds_to_regs(const PL& P, float2* line, float2 (&out)[M / kWave],
                                                int lane)
{
    float2 unused[M / kWave];
    P.template run<DIR, false, true>(line, unused, out, lane);
}

// registers -> registers
template <int M, int DIR, class PL>
__device__ __forceinline__ void fft_regs_to_regs(const PL& P, float2* line,
                                                 const float2 (&v)[M / kWave],
                                                 float2 (&out)[M / kWave], int lane)
{
    P.template run<DIR, true, true>(line, v, out, lane);
}

}  // namespace srk
