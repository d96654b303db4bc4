// srk_engine.hip -- overlap-and-save tile engine for the SRK shrinkage operator
// family on gfx950.  See srk_engine.hpp for the pass structure.
//
// Reference behaviour restated (file:line in the reference tree):
//   forward      srk_forward.cuh:354-501        (true convolution, dx^2/M^2 scale,
//                                                valid-region crop, central differences)
//   lambda_f     adjoint/srk_adjoint.cuh:25-75   (transpose of stencil + combine)
//   lambda_g     adjoint/main_adjoint.cu:206-246 (correlation = product with conj(Ghat))
//   tangent      tan_tanadj/srk_tangent.cuh:28-151
//   tangent-adj  tan_tanadj/srk_tangent_adjoint.cuh:100-159
#include "srk_engine.hpp"
#include "srk_fft.hpp"
#include "srk_fft_mfma.hpp"

#include <algorithm>
#include <type_traits>

namespace srk {

namespace {

template <class PL>
struct IsMfmaPlan { static constexpr bool value = false; };
template <int M>
struct IsMfmaPlan<FftPlanMfma<M, true>> { static constexpr bool value = true; };
template <int M>
struct IsMfmaPlan<FftPlanMfma<M, false>> { static constexpr bool value = true; };

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
// wavefront index in the block, as a scalar (per-wave pointers stay in SGPRs)
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x / kWave); }

// base[idx] for a wavefront-uniform base and a per-lane index whose byte
// offset fits in 32 bits: the address becomes base + zext(offset), which the
// compiler emits as a scalar-base load with a 32-bit vector offset.
template <class T>
__device__ __forceinline__ T ldu(const T* base, unsigned idx)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                       idx * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ void stu(T* base, unsigned idx, T v)
{
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + idx * (unsigned)sizeof(T)) = v;
}

// Streaming (nontemporal) load and store, for data that one pass reads or
// writes once and no other wavefront of the launch touches again: the input
// images of the column pass, the exchange planes written by P1 / P1A / P2 and
// the bands P2 reads.  They keep such data from evicting what is re-read
// through the caches (the spectra, the gradient stencils' neighbouring
// columns, the mirrored half-spectrum loads of P3).  Measured on the
// N = 512, B = 5000 fwd+adj step: 30.07 against 30.88 ms; N = 16384: 9.49
// against 9.91 ms.  Streaming loads in P3 (each half-spectrum element is
// read by two lanes) and streaming partial-line output stores were slower.
typedef float f2_vec __attribute__((ext_vector_type(2)));
template <class T>
__device__ __forceinline__ T ldn(const T* base, unsigned idx)
{
    const char* p = reinterpret_cast<const char*>(base) + idx * (unsigned)sizeof(T);
    if constexpr (sizeof(T) == 8) {
        return __builtin_bit_cast(T, __builtin_nontemporal_load(reinterpret_cast<const f2_vec*>(p)));
    } else {
        static_assert(sizeof(T) == 4, "streamed elements are 4 or 8 bytes");
        return __builtin_bit_cast(T, __builtin_nontemporal_load(reinterpret_cast<const unsigned*>(p)));
    }
}
template <class T>
__device__ __forceinline__ void stn(T* base, unsigned idx, T v)
{
    char* p = reinterpret_cast<char*>(base) + idx * (unsigned)sizeof(T);
    if constexpr (sizeof(T) == 8) {
        __builtin_nontemporal_store(__builtin_bit_cast(f2_vec, v), reinterpret_cast<f2_vec*>(p));
    } else {
        static_assert(sizeof(T) == 4, "streamed elements are 4 or 8 bytes");
        __builtin_nontemporal_store(__builtin_bit_cast(unsigned, v), reinterpret_cast<unsigned*>(p));
    }
}

// Load of an exchange-buffer element at base[idx].  COH: an agent-scope
// relaxed atomic load, which misses the CU's own L1 and is served by the
// XCD's L2 -- for elements another workgroup of the same launch wrote (the
// fused pipeline publishes them with store completion and a flag).
template <bool COH, class T>
__device__ __forceinline__ T ldc(const T* base, unsigned idx)
{
    if constexpr (!COH) {
        return ldu(base, idx);
    } else {
        const char* p = reinterpret_cast<const char*>(base) + idx * (unsigned)sizeof(T);
        if constexpr (sizeof(T) == 8) {
            const unsigned long long u = __hip_atomic_load(
                reinterpret_cast<unsigned long long*>(const_cast<char*>(p)), __ATOMIC_RELAXED,
                __HIP_MEMORY_SCOPE_AGENT);
            return __builtin_bit_cast(T, u);
        } else {
            static_assert(sizeof(T) == 4, "exchange elements are 4 or 8 bytes");
            const unsigned u = __hip_atomic_load(reinterpret_cast<unsigned*>(const_cast<char*>(p)),
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return __builtin_bit_cast(T, u);
        }
    }
}

// Element formats of the exchange planes: FP32 complex (float2, 8 bytes) or
// a bfloat16 pair (Bf16c, 4 bytes, real part in the low half), rounded to
// nearest even on store (v_cvt_pk_bf16_f32).  A plane keeps its FP32-sized
// slot in the exchange buffer, so plane p of slot s starts at the same byte
// address in either format; a bfloat16 plane uses the first half of its slot.
struct Bf16c {
    unsigned v;
};
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float2 unpack(float2 s) { return s; }
__device__ __forceinline__ float2 unpack(Bf16c s)
{
    return make_float2(__uint_as_float(s.v << 16), __uint_as_float(s.v & 0xffff0000u));
}
template <class S>
__device__ __forceinline__ S pack(float2 v);
template <>
__device__ __forceinline__ float2 pack<float2>(float2 v) { return v; }
template <>
__device__ __forceinline__ Bf16c pack<Bf16c>(float2 v)
{
    const bf16x2_t b = {(__bf16)v.x, (__bf16)v.y};
    return {__builtin_bit_cast(unsigned, b)};
}
template <class S>
__device__ __forceinline__ float2 ldx(const S* base, unsigned idx) { return unpack(ldu(base, idx)); }
// NT: a streaming store (the three-pass path) of an FP32 element; the
// fused pipeline keeps its exchange rings cached, since it wants them to stay
// in the L2.  bfloat16 elements are stored cached: their band segments are
// 32-byte pieces of a line, which streaming stores wrote out unmerged (P2
// conv 7.5 -> 10.1 ms in the bfloat16-exchange step at N = 512)
template <bool NT, class S>
__device__ __forceinline__ void stx(S* base, unsigned idx, float2 v)
{
    if constexpr (NT && sizeof(S) == 8)
        stn(base, idx, pack<S>(v));
    else
        stu(base, idx, pack<S>(v));
}
// plane `index` of an exchange buffer whose planes hold `elems` FP32 elements
template <class S>
__device__ __forceinline__ S* plane_at(float2* X, long long index, long long elems)
{
    return reinterpret_cast<S*>(X + index * elems);
}

// X2 plane layout ([x][k_y], column pitch HP = spec_pitch(M)): row k_y < M/2
// of column xi lies at xi HP + k_y, and the Nyquist rows (k_y = M/2) of the
// sixteen columns 16 g .. 16 g + 15 share the padding after the rows of
// column 16 g, at 16 g HP + M/2 + (xi - 16 g).  An FP32 column then fills 16
// whole 128-byte lines plus a sixteenth of a shared one, instead of 16 lines
// and a 17th that holds one element.
//
// Tiles of M >= 1024 keep each column's Nyquist element after its rows (at
// xi HP + M/2): their columns span 32 lines, and the grouped layout made the
// M = 1024 correlation pass slower.
template <int M>
constexpr bool x2_grouped() { return M <= 512; }

// element index of row k <= M/2 of column xi, relative to the plane
template <int M>
__device__ __forceinline__ long long x2_index(int xi, int k)
{
    return k == M / 2 && x2_grouped<M>()
               ? (long long)(xi & ~15) * spec_pitch(M) + M / 2 + (xi & 15)
               : (long long)xi * spec_pitch(M) + k;
}

// A column as the P3 passes address it: its rows k < M/2 from c, its Nyquist
// element at n; the zero column (spec_pitch(M) zeros) is {zcol, zcol}.
template <class S>
struct XCol {
    const S* c;
    const S* n;
};
template <int M, class S>
__device__ __forceinline__ XCol<S> x2_col(const S* plane, int xi)
{
    return {plane + (long long)xi * spec_pitch(M), plane + x2_index<M>(xi, M / 2)};
}
template <class S>
__device__ __forceinline__ XCol<S> zero_col(const float2* zcol)
{
    return {reinterpret_cast<const S*>(zcol), reinterpret_cast<const S*>(zcol)};
}

// Lane (64 - lane) mod 64's value of v (a cross-lane read through the LDS
// crossbar; no LDS memory is used).
__device__ __forceinline__ float2 mirror_lane(float2 v, int addr)
{
    return make_float2(__int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v.x))),
                       __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v.y))));
}
__device__ __forceinline__ Bf16c mirror_lane(Bf16c v, int addr)
{
    return {(unsigned)__builtin_amdgcn_ds_bpermute(addr, (int)v.v)};
}

// Two Hermitian half spectra P, Q (X2 columns, or the zero column) as loaded
// for the lane's transform inputs k = lane + 64 i (raw, so that a load can be
// issued one transform ahead), and their packing Z = P + iQ.  The DC and
// Nyquist bins are taken as real, as a real inverse transform does.
template <int M, class S, bool COH = false>
struct PairRaw {
    static constexpr int V = M / kWave;
    S p[V], q[V];

    __device__ __forceinline__ void load(const XCol<S>& P, const XCol<S>& Q, int lane)
    {
        constexpr int IN = M / 2 / kWave;  // k = M/2 at i = IN, lane 0
#pragma unroll
        for (int i = 0; i < V; ++i) {
            const int k = lane + kWave * i;
            // rows k < M/2 directly, k > M/2 by Hermitian symmetry from M - k;
            // lane 0 of i = IN (k = M/2) reads row M/2 - 1 with lane 1 (same
            // line) and then its Nyquist element
            constexpr bool G = x2_grouped<M>();
            const unsigned kk = i < IN ? k : (G && i == IN && lane == 0 ? M / 2 - 1 : M - k);
            p[i] = ldc<COH>(P.c, kk);
            q[i] = ldc<COH>(Q.c, kk);
            if (G && i == IN && lane == 0) {
                p[i] = ldc<COH>(P.n, 0u);
                q[i] = ldc<COH>(Q.n, 0u);
            }
        }
    }
    __device__ __forceinline__ void pack(float2 (&v)[V], int lane) const
    {
#pragma unroll
        for (int i = 0; i < V; ++i) {
            const int k = lane + kWave * i;
            float2 pp = unpack(p[i]), qq = unpack(q[i]);
            if (k == 0 || k == M / 2) {
                pp.y = 0.f;
                qq.y = 0.f;
            }
            v[i] = k <= M / 2 ? make_float2(pp.x - qq.y, pp.y + qq.x)
                              : make_float2(pp.x + qq.y, qq.x - pp.y);
        }
    }
};

// As PairRaw, with half the loads and registers: only the rows
// k < M/2 (lane + 64 i, i < M/128) and the Nyquist row are loaded; the
// transform inputs k > M/2 take P[M - k] from the lane that loaded it:
// M - k = (64 - lane) + 64 (V - 1 - i) for lane > 0, and 64 (V - i) (lane 0's
// own row, or the Nyquist row for i = M/128) for lane 0.  The DC and Nyquist
// bins are taken as real, as a real inverse transform does.
template <int M, class S, bool COH = false>
struct PairHalf {
    static constexpr int V = M / kWave;
    static constexpr int IN = M / 2 / kWave;  // k = M/2 at i = IN, lane 0
    S p[IN], q[IN];
    S pn, qn;  // Nyquist rows

    __device__ __forceinline__ void load(const XCol<S>& P, const XCol<S>& Q, int lane)
    {
#pragma unroll
        for (int i = 0; i < IN; ++i) {
            p[i] = ldc<COH>(P.c, (unsigned)(lane + kWave * i));
            q[i] = ldc<COH>(Q.c, (unsigned)(lane + kWave * i));
        }
        pn = ldc<COH>(P.n, 0u);
        qn = ldc<COH>(Q.n, 0u);
    }
    __device__ __forceinline__ void pack(float2 (&v)[V], int lane) const
    {
        const int addr = ((kWave - lane) & (kWave - 1)) * 4;
#pragma unroll
        for (int i = 0; i < IN; ++i) {
            float2 pp = unpack(p[i]), qq = unpack(q[i]);
            if (i == 0 && lane == 0) {  // DC
                pp.y = 0.f;
                qq.y = 0.f;
            }
            v[i] = make_float2(pp.x - qq.y, pp.y + qq.x);
        }
#pragma unroll
        for (int i = IN; i < V; ++i) {
            float2 pp = unpack(mirror_lane(p[V - 1 - i], addr));
            float2 qq = unpack(mirror_lane(q[V - 1 - i], addr));
            if (lane == 0) {
                pp = unpack(i == IN ? pn : p[V - i]);
                qq = unpack(i == IN ? qn : q[V - i]);
                if (i == IN) {  // Nyquist
                    pp.y = 0.f;
                    qq.y = 0.f;
                }
            }
            v[i] = make_float2(pp.x + qq.y, qq.x - pp.y);
        }
    }
};

// Inverse transform of P + iQ; afterwards out[i] = p(y) + i q(y) at tile row
// y = lane + 64 i.
// HALF: PairHalf loads (the P3 adjoint: N = 512 step 29.60 -> 28.89 ms).
template <int M, class PL, class S, bool COH = false, bool HALF = false>
__device__ __forceinline__ void ifft_pair(const PL& FP, float2* line, const XCol<S>& P,
                                          const XCol<S>& Q, float2 (&out)[M / kWave], int lane)
{
    std::conditional_t<HALF, PairHalf<M, S, COH>, PairRaw<M, S, COH>> r;
    r.load(P, Q, lane);
    float2 v[M / kWave];
    r.pack(v, lane);
    fft_regs_to_regs<M, +1>(FP, line, v, out, lane);
}

// Three neighbouring columns gx-1, gx, gx+1 of an image (wavefront-uniform
// pointers) and the central-difference gradient at row yy, for points whose
// stencil lies inside the image.
struct Col3 {
    const float *l, *c, *r;
};
__device__ __forceinline__ Col3 col3(const float* img, int N, int gx)
{
    const float* c = img + (long long)gx * N;
    return {c - N, c, c + N};
}
__device__ __forceinline__ float2 grad_in(const Col3& f, unsigned yy)
{
    return make_float2(0.5f * (ldu(f.r, yy) - ldu(f.l, yy)),
                       0.5f * (ldu(f.c, yy + 1u) - ldu(f.c, yy - 1u)));
}


// ---------------------------------------------------------------------------
// P1: real FFT along y of the convolved fields; one wavefront per column pair,
// 8 columns per block.
// ---------------------------------------------------------------------------
// Column block bx (8 columns) of tile t; exchange slot `slot` of X1; LDS
// lines lds[4].  256 threads.
template <int M, class PL, int NF, class S, bool NT = true>
__device__ __forceinline__ void p1_cols_item(const EngineArgs& a, const Tile& t, int slot, int bx,
                                             float2 (*lds)[line_slots(M)])
{
    constexpr int nf = NF;
    constexpr int V = M / kWave;
    constexpr int H = M / 2 + 1;
    constexpr int HP = spec_pitch(M);
    constexpr int KV = (H + kWave - 1) / kWave;  // half-spectrum values per lane
    const int lane = lane_id(), w = wave_id();
    PL FP;
    FP.load(a.tw, lane);
    const int N = a.N;
    const int x0 = bx * 8;
    const int xa = x0 + 2 * w;
    const int ca = t.p0x + xa, cb = ca + 1;
    const bool va = ca >= 0 && ca < N, vb = cb >= 0 && cb < N;
    float2* line = lds[w];
    const int RY = 1 << a.ry_log2;
    for (int fi = 0; fi < nf; ++fi) {
        if (fi) __syncthreads();  // the previous field's stores have read the lines
        // the wave's line ends up holding P (k <= M/2) at [k] and Q at [H + k]
        if (!va && !vb) {
            for (int k = lane; k < 2 * H; k += kWave) line[k] = make_float2(0.f, 0.f);
        } else {
            const float* src = (fi == 0 ? a.g0 : a.g1) + (long long)t.b * a.img_stride;
            // clamped addresses keep every load in bounds; masks zero the padding
            const float* colA = src + (long long)min(max(ca, 0), N - 1) * N;
            const float* colB = src + (long long)min(max(cb, 0), N - 1) * N;
            float2 v[V];
            // the image loads issue at raised wavefront priority, ahead of other
            // wavefronts' transform arithmetic (N = 512 step 30.07 -> 29.97 ms
            // with the same for the P2 band loads; the P3 spectrum loads were
            // slower so: 4.14 -> 4.68 ms)
            __builtin_amdgcn_s_setprio(2);
            if (va && vb && t.p0y >= 0 && t.p0y + M <= N) {
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    const int row = t.p0y + lane + kWave * i;
                    v[i] = make_float2(ldn(colA, row), ldn(colB, row));
                }
            } else {
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    const int row = t.p0y + lane + kWave * i;
                    const bool rv = row >= 0 && row < N;
                    const int rc = min(max(row, 0), N - 1);
                    const float pa = ldn(colA, rc), pb = ldn(colB, rc);
                    v[i] = make_float2((va && rv) ? pa : 0.f, (vb && rv) ? pb : 0.f);
                }
            }
            __builtin_amdgcn_s_setprio(0);
            fft_from_regs<M, -1>(FP, line, v, lane);
            // unpack Z(k), Z(M-k) into P, Q: all reads, then all writes (one wavefront,
            // LDS operations in order)
            float2 P[KV], Q[KV];
#pragma unroll
            for (int i = 0; i < KV; ++i) {
                const int k = lane + kWave * i;
                if (k < H) {
                    const float2 z = line[lslot<M>(k)];
                    const float2 zm = line[lslot<M>((M - k) & (M - 1))];
                    P[i] = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y));
                    Q[i] = make_float2(0.5f * (z.y + zm.y), -0.5f * (z.x - zm.x));
                }
            }
#pragma unroll
            for (int i = 0; i < KV; ++i) {
                const int k = lane + kWave * i;
                if (k < H) {
                    line[k] = P[i];
                    line[H + k] = Q[i];
                }
            }
        }
        __syncthreads();
        // block-wide store of columns x0 .. x0+7 into the band layout:
        // element e = (band b, column c, row r), e = b 8RY + c RY + r
        S* plane = plane_at<S>(a.X1, (long long)slot * a.np1 + fi, (long long)M * HP);
        for (int e = threadIdx.x; e < 4 * M; e += 256) {
            const int b = e >> (a.ry_log2 + 3), c = (e >> a.ry_log2) & 7, r = e & (RY - 1);
            stx<NT>(plane, ((b * M + x0) << a.ry_log2) + (e & (8 * RY - 1)),
                    lds[c >> 1][(c & 1) * H + b * RY + r]);
        }
        if (threadIdx.x < 8)  // the Nyquist row k = M/2
            stx<NT>(plane, x1_index<M>(M / 2, x0 + threadIdx.x, a.ry_log2),
                    lds[threadIdx.x >> 1][(threadIdx.x & 1) * H + M / 2]);
    }
}

template <int M, class PL, int NF, class S>
__global__ __launch_bounds__(256) void k_p1_cols(EngineArgs a)
{
    __shared__ float2 lds[4][line_slots(M)];
    p1_cols_item<M, PL, NF, S>(a, a.tiles[blockIdx.y], blockIdx.y, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
// P1A: adjoint source a_d = v0 * d_d f0 (+ v1 * d_d f1) on the valid region,
// transformed along y as a_x + i a_y; one wavefront per column, 4 per block,
// band stores staged in the wavefronts' own LDS lines as in P1.  Inside the
// valid region every stencil neighbour lies inside the image (e >= 1), so the
// loads need no bounds checks: rows outside it load at a clamped row and are
// masked.
// ---------------------------------------------------------------------------
// Column block bx (4 columns) of tile t; 256 threads.
template <int M, class PL, int NF, class S, bool NT = true>
__device__ __forceinline__ void p1_adj_src_item(const EngineArgs& a, int plane0, const Tile& t,
                                                int slot, int bx, float2 (*lds)[line_slots(M)])
{
    constexpr int nf = NF;
    constexpr int V = M / kWave;
    constexpr int HP = spec_pitch(M);
    const int lane = lane_id(), w = wave_id();
    const int N = a.N, e = a.e, W = a.W;
    // a block whose four columns all lie off the valid region has only zero
    // sources; P2 corr does not read those columns
    const int gb = t.p0x + bx * 4;
    if (gb + 3 < e || gb >= N - e) return;
    const int x = bx * 4 + w;
    const int gx = t.p0x + x;
    constexpr int H = M / 2 + 1;
    constexpr int KV = (H + kWave - 1) / kWave;
    float2* line = lds[w];
    // the wave's line ends up holding P = A_x (k <= M/2) at [k] and Q = A_y at [H + k]
    if (gx < e || gx >= N - e) {
        for (int k = lane; k < 2 * H; k += kWave) line[k] = make_float2(0.f, 0.f);
    } else {
        float2 v[V];
#pragma unroll
        for (int i = 0; i < V; ++i) v[i] = make_float2(0.f, 0.f);
        for (int fi = 0; fi < nf; ++fi) {
            // column pointers are wavefront-uniform; row offsets are unsigned 32-bit
            const float* fl = (fi == 0 ? a.f0 : a.f1) + (long long)t.b * a.img_stride +
                              (long long)(gx - 1) * N;
            const float* fc = fl + N;
            const float* fr = fc + N;
            const float* vc = (fi == 0 ? a.v0 : a.v1) + (long long)t.b * a.w_stride +
                              (long long)(gx - e) * W - e;
#pragma unroll
            for (int i = 0; i < V; ++i) {
                const int gy = t.p0y + lane + kWave * i;
                const bool ok = gy >= e && gy < N - e;
                const unsigned yy = ok ? gy : e;
                const float c = ldu(vc, yy);
                const float dxf = 0.5f * (ldu(fr, yy) - ldu(fl, yy));
                const float dyf = 0.5f * (ldu(fc, yy + 1u) - ldu(fc, yy - 1u));
                v[i].x += ok ? c * dxf : 0.f;
                v[i].y += ok ? c * dyf : 0.f;
            }
        }
        PL FP;
        FP.load(a.tw, lane);
        fft_from_regs<M, -1>(FP, line, v, lane);
        float2 P[KV], Q[KV];
#pragma unroll
        for (int i = 0; i < KV; ++i) {
            const int k = lane + kWave * i;
            if (k < H) {
                const float2 z = line[lslot<M>(k)];
                const float2 zm = line[lslot<M>((M - k) & (M - 1))];
                P[i] = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y));
                Q[i] = make_float2(0.5f * (z.y + zm.y), -0.5f * (z.x - zm.x));
            }
        }
#pragma unroll
        for (int i = 0; i < KV; ++i) {
            const int k = lane + kWave * i;
            if (k < H) {
                line[k] = P[i];
                line[H + k] = Q[i];
            }
        }
    }
    __syncthreads();
    // block-wide store of columns x0 .. x0+3 of both planes into the band layout:
    // element e = (band b, column c, row r), e = b 4RY + c RY + r
    const int x0 = bx * 4;
    const int RY = 1 << a.ry_log2;
#pragma unroll
    for (int pl = 0; pl < 2; ++pl) {
        S* plane = plane_at<S>(a.X1, (long long)slot * a.np1 + plane0 + pl, (long long)M * HP);
        for (int e2 = threadIdx.x; e2 < 2 * M; e2 += 256) {
            const int b = e2 >> (a.ry_log2 + 2), c = (e2 >> a.ry_log2) & 3, r = e2 & (RY - 1);
            stx<NT>(plane, ((b * M + x0) << a.ry_log2) + (e2 & (4 * RY - 1)),
                    lds[c][pl * H + b * RY + r]);
        }
        if (threadIdx.x < 4)  // the Nyquist row k = M/2
            stx<NT>(plane, x1_index<M>(M / 2, x0 + threadIdx.x, a.ry_log2),
                    lds[threadIdx.x][pl * H + M / 2]);
    }
}

// The grid covers the column blocks [src_b0, src_b1) only: a block with no
// valid-region column returns at once, but dispatching it costs about as much
// as a working block (at N = 512, 30 of 128 blocks per tile: P1A 4.23 -> 3.34
// ms with them left out of the grid).
template <int M, class PL, int NF, class S>
__global__ __launch_bounds__(256) void k_p1_adj_src(EngineArgs a, int plane0)
{
    __shared__ float2 lds[4][line_slots(M)];
    const int b0 = a.src_b1 > a.src_b0 ? a.src_b0 : 0;
    p1_adj_src_item<M, PL, NF, S>(a, plane0, a.tiles[blockIdx.y], blockIdx.y, b0 + blockIdx.x,
                                  lds);
}

// ---------------------------------------------------------------------------
// P2 work items.  An item is one band (RY consecutive k_y rows) of one tile
// (and, for the conv pass, one field).  Full bands come first in launch
// order: block L runs on XCD L mod 8 (workgroups are dispatched to the eight
// XCDs round robin) and only takes bands b with b mod 8 = L mod 8, so every
// XCD reads only its own eighth of the spectra, which stays resident in its
// L2.  Each such block takes up to ITEMS consecutive items of its XCD's
// list (the conv pass prefetches the next item's band into registers while it
// transforms the current one, and in band-major order keeps the band's
// spectrum rows in registers).  The Nyquist rows (k_y = M/2, a band of one
// row) follow, one item per block.
// ---------------------------------------------------------------------------
// BM (band-major): a block's consecutive items are one band of consecutive
// tiles (the conv pass keeps that band's spectrum rows in registers across
// them); otherwise consecutive bands of one tile.
template <int M, int RY, int ITEMS, bool BM = false>
struct P2Items {
    static constexpr int NBF = M / (2 * RY);  // full bands per tile
    static constexpr int BPX = NBF / 8;       // full bands per tile and XCD
    static_assert(NBF % 8 == 0, "full bands must split evenly over the eight XCDs");
    static constexpr int full_blocks(int tiles, int nf)
    {
        return 8 * ((BPX * tiles * nf + ITEMS - 1) / ITEMS);
    }
    static constexpr int blocks(int tiles, int nf) { return full_blocks(tiles, nf) + tiles * nf; }

    int xcd, first, count, nf, tiles;
    bool nyquist;
    __device__ P2Items(int tiles_, int nf_) : nf(nf_), tiles(tiles_)
    {
        const int L = blockIdx.x;
        const int fb = full_blocks(tiles, nf);
        nyquist = L >= fb;
        if (!nyquist) {
            xcd = L & 7;
            first = (L >> 3) * ITEMS;
            count = min(ITEMS, BPX * tiles * nf - first);
        } else {
            xcd = 0;
            first = L - fb;
            count = 1;
        }
    }
    // item k of this block -> tile slot, band, field.  An XCD takes its bands
    // in groups of G consecutive ones (G xcd + r + 8 G m, r < G) when G divides
    // its share, so that the G band segments of each 128-byte line of an X2
    // column are written through the same L2 (G = p2_group below).
    template <int G>
    __device__ void get(int k, int& slot, int& band, int& fi) const
    {
        const int i = first + k;
        fi = i % nf;
        const int j = i / nf;
        if (nyquist) {
            slot = j;
            band = NBF;
        } else {
            slot = BM ? j % tiles : j / BPX;
            const int q = BM ? j / tiles : j % BPX;
            band = BPX % G == 0 ? G * xcd + q % G + 8 * G * (q / G) : xcd + 8 * q;
        }
    }
    __device__ int rows(int band) const { return band < NBF ? RY : 1; }
};

// Bands whose segments (RY elements of S) fill one 128-byte line of an X2 column.
template <int RY, class S>
constexpr int p2_group()
{
    return RY * (int)sizeof(S) >= 128 ? 1 : 128 / (RY * (int)sizeof(S));
}

// LDS pitch of the P2 band lines: line_slots(M) + 16 - 16/RY, so that the
// transposing accesses below (lane -> (row lane % RY, column lane / RY)) are
// conflict-free for the band writes (32 banks, 16 lanes a cycle) and leave
// the store's reads (64 banks, 32 lanes a cycle) an eighth of their lanes in
// conflict (a pitch of line_slots + 32/RY did the opposite: the writes of
// every 16-lane group hit 8 banks twice).
template <int M, int RY>
constexpr int p2_pitch() { return line_slots(M) + 16 - 16 / RY; }

// Band of one X1 plane (M x RY elements, contiguous) -> registers of the
// block's threads (fetch_band) -> LDS lines (put_band): line r holds k_y row
// k0 + r along x.  Rows >= nrows are zero.
template <int M, int RY, int NW, class S>
struct BandRegs {
    static constexpr int T = NW * kWave, IT = M * RY / T;
    static_assert((T / RY) % 8 == 0 && (M * RY) % T == 0, "band load shape");
    S v[IT];  // as stored: a bfloat16 prefetch holds half the registers

    // columns outside [xlo, xhi) (tile-local) are taken as zero and not read;
    // COH: the band was written by another workgroup of the same launch
    template <bool COH = false>
    __device__ __forceinline__ void fetch(const S* __restrict__ band, int nrows, int xlo = 0,
                                          int xhi = M)
    {
        constexpr int XS = T / RY;
        const bool ok = (int)(threadIdx.x % RY) < nrows;
        const int x0 = threadIdx.x / RY;
        // a full band is x-major with RY rows per x; the Nyquist band (one
        // row) is contiguous along x
        const bool nyq = nrows == 1;
        const S* src = band + (nyq ? x0 : (int)threadIdx.x);
        const int step = nyq ? XS : T;
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int x = x0 + i * XS;
            v[i] = pack<S>(make_float2(0.f, 0.f));
            if (ok && x >= xlo && x < xhi)
                v[i] = COH ? ldc<true>(src, i * step) : ldn(src, i * step);
        }
    }
    __device__ __forceinline__ void put(float2* L) const
    {
        constexpr int XS = T / RY, LP = p2_pitch<M, RY>();
        float2* d = L + (threadIdx.x % RY) * LP + lslot<M>(threadIdx.x / RY);
#pragma unroll
        for (int i = 0; i < IT; ++i) d[i * slot_off<M>(XS)] = unpack(v[i]);
    }
};

// LDS lines -> X2 plane (layout at x2_col): X2 column xi holds tile-local
// column lo + xi; columns xi in [xs, xe) are stored, those outside [0, M) as
// zero.
template <int M, int RY, int NW, class S, bool NT = true>
__device__ __forceinline__ void store_band(S* __restrict__ plane, const float2* L, int k0,
                                           int nrows, int lo, int xs, int xe)
{
    constexpr int T = NW * kWave, XS = T / RY, HP = spec_pitch(M), LP = p2_pitch<M, RY>();
    const int r = threadIdx.x % RY, xi0 = xs + threadIdx.x / RY;
    if (r >= nrows) return;
    const float2* src = L + r * LP;
    if (x2_grouped<M>() && k0 == M / 2) {
        // the Nyquist row: sixteen columns' elements fill one line, written by
        // several wavefronts, so the stores stay cached (merged in the L2)
        for (int xi = xi0; xi < xe; xi += XS) {
            const int x = lo + xi;
            float2 v = make_float2(0.f, 0.f);
            if (x >= 0 && x < M) v = src[lslot<M>(x)];
            plane[x2_index<M>(xi, M / 2)] = pack<S>(v);
        }
        return;
    }
    S* d = plane + (long long)xi0 * HP + k0 + r;
    for (int xi = xi0; xi < xe; xi += XS, d += (long long)XS * HP) {
        const int x = lo + xi;
        float2 v = make_float2(0.f, 0.f);
        if (x >= 0 && x < M) v = src[lslot<M>(x)];
        if constexpr (NT && sizeof(S) == 8)  // FP32 only, as stx
            stn(d, 0u, pack<S>(v));
        else
            *d = pack<S>(v);
    }
}

// ---------------------------------------------------------------------------
// P2 conv: x-FFT of each convolved-field band row, products with Gx and Gy in
// registers, inverse x-FFTs, crop; field fi -> X2 planes 2fi (u_x), 2fi+1 (u_y).
// ---------------------------------------------------------------------------
// Large tiles (vector-ALU plan, M >= 1024) keep one LDS line per band row
// instead of two, so that two blocks fit a CU (74 instead of 148 KB at both
// M = 1024 and 2048 with their band heights): the corr pass transforms its two
// source planes one after the other; the conv pass transforms and stores u_x
// first and then u_y through the same lines, keeping the u_y spectrum in
// registers meanwhile, with one item per block (no prefetch registers).  At
// M = 1024 (N = 16384 step) the one-line corr pass took 2.33-2.43 ms against
// 2.80 for two lines; the conv pass measured the same either way (4.45-4.58 ms).
// The corr pass keeps one line per row at M = 512 as well, at six wavefronts
// per SIMD (80 registers, three 38 KB blocks per CU): 4.10 -> 3.87 ms in the
// N = 512, B = 5000 step; the conv pass there keeps two lines and its
// four-item blocks.
enum P2Pass { kP2ConvPass = 0, kP2CorrPass = 1 };
template <int M, class PL, int PASS>
constexpr bool p2_one_line()
{
    return !IsMfmaPlan<PL>::value && (M >= 1024 || (PASS == kP2CorrPass && M == 512));
}
// Items per conv block: with two lines per row (M <= 512) the FP32 plans take
// four items of one band (spectrum rows loaded once; N = 512 step 28.73 ->
// 28.58 ms, where two items per block, band or tile major, tie) when the
// planes are FP32; bfloat16 planes keep two tile-major items, so that the
// four bands whose 32-byte segments fill an X2 line are written together
// (band-major, bf16-exchange P2 conv 7.3 -> 9.8 ms)
template <int M, class PL, class S>
constexpr bool p2c_keep()
{
    return !p2_one_line<M, PL, kP2ConvPass>() && !IsMfmaPlan<PL>::value && sizeof(S) == 8;
}
template <int M, class PL, class S>
constexpr int p2c_items()
{
    return p2_one_line<M, PL, kP2ConvPass>() ? 1 : p2c_keep<M, PL, S>() ? 4 : 2;
}

// X2 columns [xs, xe) of a tile's u planes that P3 reads: u enters the
// operators only multiplied by a field that vanishes off the valid region
// (grad f in the forward, the cotangent in the adjoints), so only the global
// columns [e, N - e) are stored.  X2 column xi is global column
// ox0 - halo2 + xi.
__device__ __forceinline__ void u_columns(const EngineArgs& a, const Tile& t, int& xs, int& xe)
{
    const int cnt = t.ox1 - t.ox0 + 2 * a.halo2;
    xs = max(0, a.e - t.ox0 + a.halo2);
    xe = min(cnt, a.N - a.e - t.ox0 + a.halo2);
}

// two 8-wavefront blocks per CU at M = 1024 need at most 128 registers
template <int M, class PL, int PASS>
constexpr int p2_waves_per_eu()
{
    return !p2_one_line<M, PL, PASS>() ? 1 : M == 1024 ? 4 : M == 512 ? 6 : 1;
}

template <int M, int RY, int NW, class PL, class S>
__global__ __launch_bounds__(NW * kWave)
__attribute__((amdgpu_waves_per_eu(p2_waves_per_eu<M, PL, kP2ConvPass>())))
void k_p2_conv(EngineArgs a, int nf, int tiles)
{
    constexpr int V = M / kWave;
    constexpr int H = M / 2 + 1;
    constexpr int HP = spec_pitch(M);
    constexpr int LP = p2_pitch<M, RY>();
    constexpr bool ONE = p2_one_line<M, PL, kP2ConvPass>();
    static_assert(!ONE || NW == RY, "one band row per wavefront");
    // with one band row per wavefront and two lines per row, the row's
    // spectra stay in registers while the block's items share the band
    constexpr bool KEEP = p2c_keep<M, PL, S>() && NW == RY;
    extern __shared__ float2 sm[];
    float2* A = sm;
    float2* Bs = ONE ? sm : sm + RY * LP;
    const int lane = lane_id(), w = wave_id();
    PL FP;
    FP.load(a.tw, lane);
    const P2Items<M, RY, p2c_items<M, PL, S>(), KEEP> items(tiles, nf);
    float2 gxv[V], gyv[V];
    int gband = -1;  // band whose spectrum rows gxv, gyv hold (KEEP)
    auto band_src = [&](int slot, int band, int fi) {
        return plane_at<const S>(a.X1, (long long)slot * a.np1 + fi, (long long)M * HP) +
               (long long)band * RY * M;
    };
    BandRegs<M, RY, NW, S> pf;
    int slot, band, fi;
    items.template get<p2_group<RY, S>()>(0, slot, band, fi);
    __builtin_amdgcn_s_setprio(2);  // band loads issue first, as the P1 image loads
    pf.fetch(band_src(slot, band, fi), items.rows(band));
    __builtin_amdgcn_s_setprio(0);
    for (int k = 0; k < items.count; ++k) {
        items.template get<p2_group<RY, S>()>(k, slot, band, fi);
        const int k0 = band * RY;
        const int nrows = items.rows(band);
        if (k) __syncthreads();  // the previous item's stores have read A and Bs
        pf.put(A);
        __syncthreads();
        if (k + 1 < items.count) {
            int s2, b2, f2;
            items.template get<p2_group<RY, S>()>(k + 1, s2, b2, f2);
            __builtin_amdgcn_s_setprio(2);
            pf.fetch(band_src(s2, b2, f2), items.rows(b2));
            __builtin_amdgcn_s_setprio(0);
        }
        const long long pe = (long long)a.S2 * HP;
        if constexpr (ONE) {
            const Tile t = a.tiles[slot];
            const int lo = t.ox0 - t.p0x - a.halo2;
            int xs, xe;
            u_columns(a, t, xs, xe);
            const bool row = w < nrows;
            float2* la = A + w * LP;
            float2 U[V];
            if (row) {
                const float2* gxr = a.GT + (long long)(k0 + w) * M + lane;
                const float2* gyr = a.GT + (long long)(H + k0 + w) * M + lane;
                float2 X[V];
                // the Gy loads are in flight during the forward transform (the
                // Gx row as well would spill under the 128-register bound)
#pragma unroll
                for (int i = 0; i < V; ++i) U[i] = ldu(gyr, kWave * i);
                fft_lds_to_regs<M, -1>(FP, la, X, lane);
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    U[i] = cmul(X[i], U[i]);
                    X[i] = cmul(X[i], ldu(gxr, kWave * i));
                }
                fft_from_regs<M, +1>(FP, la, X, lane);
            }
            __syncthreads();
            store_band<M, RY, NW>(plane_at<S>(a.X2, (long long)slot * a.np2 + 2 * fi, pe), A, k0,
                                  nrows, lo, xs, xe);
            __syncthreads();  // every row of u_x is stored: the lines take u_y
            if (row) fft_from_regs<M, +1>(FP, la, U, lane);
            __syncthreads();
            store_band<M, RY, NW>(plane_at<S>(a.X2, (long long)slot * a.np2 + 2 * fi + 1, pe), A,
                                  k0, nrows, lo, xs, xe);
            continue;
        }
        for (int r = w; r < nrows; r += NW) {
            float2* la = A + r * LP;
            float2* lb = Bs + r * LP;
            const float2* gxr = a.GT + (long long)(k0 + r) * M + lane;
            const float2* gyr = a.GT + (long long)(H + k0 + r) * M + lane;
            if (!KEEP || band != gband) {
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    gxv[i] = ldu(gxr, kWave * i);
                    gyv[i] = ldu(gyr, kWave * i);
                }
                gband = band;
            }
            float2 X[V], U[V];
            fft_lds_to_regs<M, -1>(FP, la, X, lane);
#pragma unroll
            for (int i = 0; i < V; ++i) {
                U[i] = cmul(X[i], gyv[i]);
                X[i] = cmul(X[i], gxv[i]);
            }
            fft_from_regs<M, +1>(FP, la, X, lane);
            fft_from_regs<M, +1>(FP, lb, U, lane);
        }
        __syncthreads();
        const Tile t = a.tiles[slot];
        const int lo = t.ox0 - t.p0x - a.halo2;
        int xs, xe;
        u_columns(a, t, xs, xe);
        store_band<M, RY, NW>(plane_at<S>(a.X2, (long long)slot * a.np2 + 2 * fi, pe), A, k0,
                              nrows, lo, xs, xe);
        store_band<M, RY, NW>(plane_at<S>(a.X2, (long long)slot * a.np2 + 2 * fi + 1, pe), Bs,
                              k0, nrows, lo, xs, xe);
    }
}

// ---------------------------------------------------------------------------
// P2 corr: x-FFT of the adjoint sources, L = conj(Gx) A_x + conj(Gy) A_y in
// registers, inverse x-FFT, crop -> X2 plane planeOut.
// ---------------------------------------------------------------------------
template <int M, int RY, int NW, class PL, class S>
__global__ __launch_bounds__(NW * kWave)
__attribute__((amdgpu_waves_per_eu(p2_waves_per_eu<M, PL, kP2CorrPass>())))
void k_p2_corr(EngineArgs a, int planeIn, int planeOut, int tiles)
{
    constexpr int V = M / kWave;
    constexpr int H = M / 2 + 1;
    constexpr int HP = spec_pitch(M);
    constexpr int LP = p2_pitch<M, RY>();
    constexpr bool ONE = p2_one_line<M, PL, kP2CorrPass>();
    static_assert(!ONE || NW == RY, "one band row per wavefront");
    extern __shared__ float2 sm[];
    float2* A = sm;
    float2* Bs = ONE ? sm : sm + RY * LP;
    const int lane = lane_id(), w = wave_id();
    PL FP;
    FP.load(a.tw, lane);
    const P2Items<M, RY, 1> items(tiles, 1);
    auto band_src = [&](int slot, int band, int p) {
        return plane_at<const S>(a.X1, (long long)slot * a.np1 + planeIn + p, (long long)M * HP) +
               (long long)band * RY * M;
    };
    // one item per block: a prefetch of two planes would cost the registers
    // of the second block per CU
    int slot, band, fi;
    for (int k = 0; k < items.count; ++k) {
        items.template get<p2_group<RY, S>()>(k, slot, band, fi);
        const int k0 = band * RY;
        const int nrows = items.rows(band);
        const Tile t = a.tiles[slot];
        const int lo = t.ox0 - t.p0x - a.halo2;
        const int cnt = t.ox1 - t.ox0 + 2 * a.halo2;
        S* dst = plane_at<S>(a.X2, (long long)slot * a.np2 + planeOut, (long long)a.S2 * HP);
        // the sources vanish off the valid region: P1A leaves columns there unwritten
        const int sx0 = a.e - t.p0x, sx1 = a.N - a.e - t.p0x;
        if constexpr (ONE) {
            const bool row = w < nrows;
            float2* la = A + w * LP;
            BandRegs<M, RY, NW, S> pb;
            {
                BandRegs<M, RY, NW, S> pa;
                pa.fetch(band_src(slot, band, 0), nrows, sx0, sx1);
                if (k) __syncthreads();  // the previous item's store has read A
                pa.put(A);
            }
            __syncthreads();
            pb.fetch(band_src(slot, band, 1), nrows, sx0, sx1);  // in flight during the first transform
            float2 X[V];
            if (row) fft_lds_to_regs<M, -1>(FP, la, X, lane);
            __syncthreads();  // every line is read: the lines take the second plane
            pb.put(A);
            // the conj(Gx) loads are in flight across the barrier, the conj(Gy)
            // loads during the second transform
            const float2* gxr = a.GT + (long long)(k0 + w) * M + lane;
            const float2* gyr = a.GT + (long long)(H + k0 + w) * M + lane;
            float2 G[V];
            if (row) {
#pragma unroll
                for (int i = 0; i < V; ++i) G[i] = ldu(gxr, kWave * i);
            }
            __syncthreads();
            if (row) {
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    X[i] = cmulc(X[i], G[i]);
                    G[i] = ldu(gyr, kWave * i);
                }
                float2 Y[V];
                fft_lds_to_regs<M, -1>(FP, la, Y, lane);
#pragma unroll
                for (int i = 0; i < V; ++i) X[i] = cadd(X[i], cmulc(Y[i], G[i]));
                fft_from_regs<M, +1>(FP, la, X, lane);
            }
            __syncthreads();
            store_band<M, RY, NW>(dst, A, k0, nrows, lo, 0, cnt);
            continue;
        }
        {
            BandRegs<M, RY, NW, S> pa, pb;
            pa.fetch(band_src(slot, band, 0), nrows, sx0, sx1);
            pb.fetch(band_src(slot, band, 1), nrows, sx0, sx1);
            if (k) __syncthreads();  // the previous item's store has read A
            pa.put(A);
            pb.put(Bs);
        }
        __syncthreads();
        for (int r = w; r < nrows; r += NW) {
            float2* la = A + r * LP;
            float2* lb = Bs + r * LP;
            const float2* gxr = a.GT + (long long)(k0 + r) * M + lane;
            const float2* gyr = a.GT + (long long)(H + k0 + r) * M + lane;
            float2 gxv[V], gyv[V];
#pragma unroll
            for (int i = 0; i < V; ++i) {
                gxv[i] = ldu(gxr, kWave * i);
                gyv[i] = ldu(gyr, kWave * i);
            }
            float2 X[V], Y[V];
            fft_lds_to_regs<M, -1>(FP, la, X, lane);
            fft_lds_to_regs<M, -1>(FP, lb, Y, lane);
#pragma unroll
            for (int i = 0; i < V; ++i) X[i] = cadd(cmulc(X[i], gxv[i]), cmulc(Y[i], gyv[i]));
            fft_from_regs<M, +1>(FP, la, X, lane);
        }
        __syncthreads();
        store_band<M, RY, NW>(dst, A, k0, nrows, lo, 0, cnt);
    }
}

// ---------------------------------------------------------------------------
// Row passes of the fused pipeline (k_fused below): one band of kFusedRY k_y
// rows of one tile per call, kFusedNW wavefronts with kFusedRY / kFusedNW rows
// each and one LDS line per row (kFusedRY lines of pitch p2_pitch, 37.8 KB at
// M = 512), so that four 256-thread workgroups share a CU.  The per-row
// arithmetic is that of k_p2_conv / k_p2_corr; X1 bands are read with
// coherent loads (written by other workgroups of the launch).
// ---------------------------------------------------------------------------
constexpr int kFusedRY = 8, kFusedNW = 4;

template <int M>
constexpr int fused_p2_floats() { return kFusedRY * p2_pitch<M, kFusedRY>() * 2; }

// u_x, u_y of band `band` of field 0 (X1 plane 0 -> X2 planes 0, 1)
template <int M, class PL, class S>
__device__ __forceinline__ void p2_conv_band_fused(const EngineArgs& a, const Tile& t, int slot,
                                                   int band, float2* A)
{
    constexpr int RY = kFusedRY, NW = kFusedNW, RPW = RY / NW;
    constexpr int V = M / kWave, H = M / 2 + 1, HP = spec_pitch(M), LP = p2_pitch<M, RY>();
    constexpr int NBF = M / (2 * RY);
    const int lane = lane_id(), w = wave_id();
    PL FP;
    FP.load(a.tw, lane);
    const int k0 = band * RY, nrows = band < NBF ? RY : 1;
    {
        BandRegs<M, RY, NW, S> pf;
        pf.template fetch<true>(
            plane_at<const S>(a.X1, (long long)slot * a.np1, (long long)M * HP) +
                (long long)band * RY * M,
            nrows);
        pf.put(A);
    }
    __syncthreads();
    float2 U[RPW][V];
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int r = w + q * NW;
        if (r < nrows) {
            float2* la = A + r * LP;
            const float2* gxr = a.GT + (long long)(k0 + r) * M + lane;
            const float2* gyr = a.GT + (long long)(H + k0 + r) * M + lane;
#pragma unroll
            for (int i = 0; i < V; ++i) U[q][i] = ldu(gyr, kWave * i);
            float2 X[V];
            fft_lds_to_regs<M, -1>(FP, la, X, lane);
#pragma unroll
            for (int i = 0; i < V; ++i) {
                U[q][i] = cmul(X[i], U[q][i]);
                X[i] = cmul(X[i], ldu(gxr, kWave * i));
            }
            fft_from_regs<M, +1>(FP, la, X, lane);
        }
    }
    __syncthreads();
    const int lo = t.ox0 - t.p0x - a.halo2;
    int xs, xe;
    u_columns(a, t, xs, xe);
    const long long pe = (long long)a.S2 * HP;
    store_band<M, RY, NW, S, false>(plane_at<S>(a.X2, (long long)slot * a.np2, pe), A, k0, nrows,
                                    lo, xs, xe);
    __syncthreads();  // every row of u_x is stored: the lines take u_y
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int r = w + q * NW;
        if (r < nrows) fft_from_regs<M, +1>(FP, A + r * LP, U[q], lane);
    }
    __syncthreads();
    store_band<M, RY, NW, S, false>(plane_at<S>(a.X2, (long long)slot * a.np2 + 1, pe), A, k0,
                                    nrows, lo, xs, xe);
}

// conj(Gx) A_x + conj(Gy) A_y of band `band` (X1 planes planeIn, planeIn + 1
// -> X2 plane planeOut)
template <int M, class PL, class S>
__device__ __forceinline__ void p2_corr_band_fused(const EngineArgs& a, const Tile& t, int slot,
                                                   int band, int planeIn, int planeOut, float2* A)
{
    constexpr int RY = kFusedRY, NW = kFusedNW, RPW = RY / NW;
    constexpr int V = M / kWave, H = M / 2 + 1, HP = spec_pitch(M), LP = p2_pitch<M, RY>();
    constexpr int NBF = M / (2 * RY);
    const int lane = lane_id(), w = wave_id();
    PL FP;
    FP.load(a.tw, lane);
    const int k0 = band * RY, nrows = band < NBF ? RY : 1;
    auto band_src = [&](int p) {
        return plane_at<const S>(a.X1, (long long)slot * a.np1 + planeIn + p, (long long)M * HP) +
               (long long)band * RY * M;
    };
    // the sources vanish off the valid region: P1A leaves columns there unwritten
    const int sx0 = a.e - t.p0x, sx1 = a.N - a.e - t.p0x;
    BandRegs<M, RY, NW, S> pb;
    {
        BandRegs<M, RY, NW, S> pa;
        pa.template fetch<true>(band_src(0), nrows, sx0, sx1);
        pa.put(A);
    }
    __syncthreads();
    pb.template fetch<true>(band_src(1), nrows, sx0, sx1);  // in flight during the transforms
    float2 X[RPW][V];
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int r = w + q * NW;
        if (r < nrows) fft_lds_to_regs<M, -1>(FP, A + r * LP, X[q], lane);
    }
    __syncthreads();  // every line is read: the lines take the second plane
    pb.put(A);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPW; ++q) {
        const int r = w + q * NW;
        if (r < nrows) {
            const float2* gxr = a.GT + (long long)(k0 + r) * M + lane;
            const float2* gyr = a.GT + (long long)(H + k0 + r) * M + lane;
            float2 G[V], Y[V];
#pragma unroll
            for (int i = 0; i < V; ++i) G[i] = ldu(gyr, kWave * i);
            fft_lds_to_regs<M, -1>(FP, A + r * LP, Y, lane);
#pragma unroll
            for (int i = 0; i < V; ++i)
                X[q][i] = cadd(cmulc(X[q][i], ldu(gxr, kWave * i)), cmulc(Y[i], G[i]));
            fft_from_regs<M, +1>(FP, A + r * LP, X[q], lane);
        }
    }
    __syncthreads();
    const int lo = t.ox0 - t.p0x - a.halo2;
    const int cnt = t.ox1 - t.ox0 + 2 * a.halo2;
    store_band<M, RY, NW, S, false>(plane_at<S>(a.X2, (long long)slot * a.np2 + planeOut,
                                                (long long)a.S2 * HP),
                                    A, k0, nrows, lo, 0, cnt);
}

// ---------------------------------------------------------------------------
// P3 forward / tangent: inverse y-FFT of (u_x + i u_y) per output column,
// gradient stencil and combine, valid-region store.
// ---------------------------------------------------------------------------
// Output columns 4 bx .. 4 bx + 3 of tile t (one per wavefront); 256 threads.
// COH: the X2 columns were written by another workgroup of the same launch.
template <int M, class PL, int NF, class S, bool COH = false>
__device__ __forceinline__ void p3_fwd_item(const EngineArgs& a, const Tile& t, int slot, int bx,
                                            float2 (*lds)[line_slots(M)])
{
    constexpr int V = M / kWave;
    constexpr int HP = spec_pitch(M);
    const int lane = lane_id(), w = wave_id();
    const int xi = bx * 4 + w;
    if (xi >= t.ox1 - t.ox0) return;
    PL FP;
    FP.load(a.tw, lane);
    const int N = a.N, e = a.e, W = a.W;
    const int gx = t.ox0 + xi;
    float2* line = lds[w];
    // this column of X2 plane p
    auto col = [&](int p) {
        return x2_col<M>(plane_at<const S>(a.X2, (long long)slot * a.np2 + p, (long long)a.S2 * HP),
                         xi);
    };
    // output rows [oy0, oy1) and the column lie inside the valid region, so
    // every gradient stencil is inside the image
    const Col3 F0 = col3(a.f0 + (long long)t.b * a.img_stride, N, gx);
    const long long ob = (long long)t.b * a.w_stride + (long long)(gx - e) * W - e;
    float* out = a.out0 + ob;  // indexed by the global row

    // u(g) for this column at tile rows y = lane + 64 j
    float2 u0[V];
    ifft_pair<M, PL, S, COH>(FP, line, col(0), col(1), u0, lane);
    if constexpr (NF == 1) {
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const int gy = t.p0y + lane + kWave * j;
            if (gy >= t.oy0 && gy < t.oy1) {  // conditional: keeps the loads late
                const float2 g0 = grad_in(F0, (unsigned)gy);
                stu(out, (unsigned)gy, u0[j].x * g0.x + u0[j].y * g0.y);
            }
        }
        if (a.ux_out) {
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int gy = t.p0y + lane + kWave * j;
                if (gy >= t.oy0 && gy < t.oy1) {
                    stu(a.ux_out + ob, (unsigned)gy, u0[j].x);
                    stu(a.uy_out + ob, (unsigned)gy, u0[j].y);
                }
            }
        }
    } else {
        // tangent: + u(dg) . grad f + u(g) . grad df
        const Col3 F1 = col3(a.f1 + (long long)t.b * a.img_stride, N, gx);
        float2 u1[V];
        ifft_pair<M, PL, S, COH>(FP, line, col(2), col(3), u1, lane);
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const int gy = t.p0y + lane + kWave * j;
            if (gy >= t.oy0 && gy < t.oy1) {
                const float2 g0 = grad_in(F0, (unsigned)gy), g1 = grad_in(F1, (unsigned)gy);
                stu(out, (unsigned)gy,
                    u1[j].x * g0.x + u1[j].y * g0.y + u0[j].x * g1.x + u0[j].y * g1.y);
            }
        }
    }
}

// P3 forward with one field, as p3_fwd_item, but with the image columns of
// the gradient stencils shared through LDS.  After its transform a
// wavefront's line is free: it takes the wavefront's own image column at
// every tile row, neighbours read it there instead of loading columns gx - 1
// and gx + 1 again, and the y-differences come from the same line; only the
// block's edge wavefronts load an outer neighbour column from memory.  The
// valid rows of a tile lie at least one row inside it, so y - 1 and y + 1 are
// tile rows.  Same arithmetic as p3_fwd_item: bit-identical results.
template <int M, class PL, class S>
__device__ __forceinline__ void p3_fwd_shared_item(const EngineArgs& a, const Tile& t, int slot,
                                                   int bx, float2 (*lds)[line_slots(M)])
{
    constexpr int V = M / kWave;
    constexpr int HP = spec_pitch(M);
    const int lane = lane_id(), w = wave_id();
    const int cnt = t.ox1 - t.ox0;
    const int xi = bx * 4 + w;
    const bool active = xi < cnt;  // the active wavefronts of a block are a prefix
    const int N = a.N, e = a.e, W = a.W;
    float* fline = reinterpret_cast<float*>(lds[w]);
    // image column and output addresses: formed after the spectrum loads at
    // M <= 512 (63 instead of 67 registers, eight wavefronts per SIMD), before
    // them at M = 1024 (115 instead of 142 registers)
    const int gx = t.ox0 + xi;
    const float* fc = nullptr;
    long long ob = 0;
    auto addresses = [&] {
        fc = a.f0 + (long long)t.b * a.img_stride + (long long)gx * N;
        ob = (long long)t.b * a.w_stride + (long long)(gx - e) * W - e;
    };
    if constexpr (M > 512) addresses();
    // M <= 512: half loads with the lane exchange (PairHalf; P3 forward 3.98
    // -> 3.78 ms at N = 512, B = 5000); M = 1024 keeps PairRaw, which
    // measured faster there
    using Raw = std::conditional_t<(M <= 512), PairHalf<M, S>, PairRaw<M, S>>;
    PL FP;
    Raw raw;
    if (active) {
        const long long pe = (long long)a.S2 * HP;
        FP.load(a.tw, lane);
        raw.load(x2_col<M>(plane_at<const S>(a.X2, (long long)slot * a.np2, pe), xi),
                 x2_col<M>(plane_at<const S>(a.X2, (long long)slot * a.np2 + 1, pe), xi), lane);
    }
    if constexpr (M <= 512) addresses();
    float2 u0[M / kWave];
    if (active) {
        // the image column's loads go out once the spectrum values are packed
        // (their raw registers are free then) and land during the transform;
        // rows outside the image load a clamped row, which no valid row reads
        float2 z[M / kWave];
        raw.pack(z, lane);
        float fcol[M / kWave];
#pragma unroll
        for (int j = 0; j < V; ++j)
            fcol[j] = ldu(fc, (unsigned)min(max(t.p0y + lane + kWave * j, 0), N - 1));
        fft_regs_to_regs<M, +1>(FP, lds[w], z, u0, lane);
#pragma unroll
        for (int j = 0; j < V; ++j) fline[lane + kWave * j] = fcol[j];
    }
    __syncthreads();
    if (!active) return;
    const float* fleft = w > 0 ? reinterpret_cast<const float*>(lds[w - 1]) : nullptr;
    const float* fright =
        w < 3 && xi + 1 < cnt ? reinterpret_cast<const float*>(lds[w + 1]) : nullptr;
    float* out = a.out0 + ob;  // indexed by the global row
#pragma unroll
    for (int j = 0; j < V; ++j) {
        const int y = lane + kWave * j;
        const int gy = t.p0y + y;
        if (gy >= t.oy0 && gy < t.oy1) {
            const float l = fleft ? fleft[y] : ldu(fc - N, (unsigned)gy);
            const float r = fright ? fright[y] : ldu(fc + N, (unsigned)gy);
            const float2 g = make_float2(0.5f * (r - l), 0.5f * (fline[y + 1] - fline[y - 1]));
            stu(out, (unsigned)gy, u0[j].x * g.x + u0[j].y * g.y);
        }
    }
    if (a.ux_out) {
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const int gy = t.p0y + lane + kWave * j;
            if (gy >= t.oy0 && gy < t.oy1) {
                stu(a.ux_out + ob, (unsigned)gy, u0[j].x);
                stu(a.uy_out + ob, (unsigned)gy, u0[j].y);
            }
        }
    }
}

template <int M, class PL, int NF, class S>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(M <= 512 && NF == 1 ? 8 : 1))) void k_p3_fwd(EngineArgs a)
{
    __shared__ float2 lds[4][line_slots(M)];
    if constexpr (NF == 1)
        p3_fwd_shared_item<M, PL, S>(a, a.tiles[blockIdx.y], blockIdx.y, blockIdx.x, lds);
    else
        p3_fwd_item<M, PL, NF, S>(a, a.tiles[blockIdx.y], blockIdx.y, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
// P3 adjoint.  A block owns 2 * P3A_WAVES consecutive output columns, a
// wavefront the pair (gx, gx+1):
//   lambda_f(c, r) = (w_x(c-1, r) - w_x(c+1, r))/2 + (w_y(c, r-1) - w_y(c, r+1))/2,
//   w_d = u_d(g) v0 (+ u_d(dg) v1), zero outside the valid region;
//   lambda_g = inverse y-FFT of the correlation plane.
// Each wavefront transforms u_x + i u_y once per own column; w_x goes to a
// block array in LDS so that neighbours read it instead of transforming u_x
// again, and the last wavefront also transforms the two halo columns of the
// block (u_x of the columns left and right of it, packed in one transform).
// X2 column index of global column c is c - ox0 + 1; cnt + 2 columns are
// stored, and columns beyond are zero.
// ---------------------------------------------------------------------------
constexpr int P3A_WAVES = 4;

// w_x of the 2 P3A_WAVES + 2 columns of a P3 adjoint block, rows y
template <int M>
constexpr int p3a_wx_floats() { return (2 * P3A_WAVES + 2) * M; }

// Output columns CBK bx .. CBK bx + CBK - 1 of tile t; P3A_WAVES wavefronts.
// lds: one line per wavefront; wx: p3a_wx_floats<M>() floats.  COH as p3_fwd_item.
template <int M, class PL, int NF, class SL, bool COH = false>
__device__ __forceinline__ void p3_adj_item(const EngineArgs& a, int with_u, int planeL,
                                            const Tile& t, int slot, int bx,
                                            float2 (*lds)[line_slots(M)], float* wx)
{
    constexpr int nf = NF;
    constexpr int V = M / kWave;
    constexpr int HP = spec_pitch(M);
    constexpr int CBK = 2 * P3A_WAVES;  // output columns per block
    const int lane = lane_id(), w = wave_id();
    const int jb = bx * CBK;
    const int cnt = t.ox1 - t.ox0;
    if (jb >= cnt) return;  // whole block beyond the tile's columns
    const int j0 = jb + 2 * w;
    PL FP;
    FP.load(a.tw, lane);
    const int N = a.N, e = a.e, W = a.W;
    const int gx = t.ox0 + j0;
    float2* line = lds[w];
    const long long pstride = (long long)a.S2 * HP;
    const int ncols = cnt + 2;
    // X2 column spectrum of output column j (tile-local), or the zero column
    // beyond the stored ones; the u planes are FP32, the lambda_g plane SL
    auto colS = [&](auto fmt, int plane, int j) {
        using T = decltype(fmt);
        const int xi2 = j + 1;
        return xi2 < ncols
                   ? x2_col<M>(plane_at<const T>(a.X2, (long long)slot * a.np2 + plane, pstride),
                               xi2)
                   : zero_col<T>(a.zcol);
    };
    // u planes: stored only for global columns [e, N - e) (u_columns)
    auto col = [&](int plane, int j) {
        const int c = t.ox0 + j;
        return c >= a.e && c < a.N - a.e ? colS(float2{}, plane, j)
                                         : zero_col<float2>(a.zcol);
    };
    const long long ib = (long long)t.b * a.img_stride;

    if (with_u) {
        // lambda_f at columns gx (acc0) and gx+1 (acc1), tile rows y = lane + 64 j
        float acc0[V], acc1[V];
#pragma unroll
        for (int j = 0; j < V; ++j) acc0[j] = acc1[j] = 0.f;
        for (int fi = 0; fi < nf; ++fi) {
            const float* vv = (fi == 0 ? a.v0 : a.v1) + (long long)t.b * a.w_stride;
            // cotangent of global column c at this lane's rows, zero off the valid region
            auto cot = [&](int c, float (&cv)[V]) {
                const bool colv = c >= e && c < N - e;
                const float* vcol = vv + (long long)(colv ? c - e : 0) * W - e;
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    const int gy = t.p0y + lane + kWave * j;
                    const bool ok = colv && gy >= e && gy < N - e;
                    cv[j] = ok ? vcol[gy] : 0.f;  // conditional: keeps the loads late
                }
            };
            const int px = 2 * fi, py = 2 * fi + 1;
            float2 u[V];
            float wy0[V], cv[V];
            // own columns: u_x + i u_y, w_x to LDS, w_y kept; off the valid
            // region the cotangent is zero and the transform is skipped
            auto u_pair = [&](int j) {
                const int c = t.ox0 + j;
                if (c >= e && c < N - e) {
                    ifft_pair<M, PL, float2, COH, true>(FP, line, col(px, j), col(py, j), u, lane);
                } else {
#pragma unroll
                    for (int q = 0; q < V; ++q) u[q] = make_float2(0.f, 0.f);
                }
            };
            // both own columns' cotangents load before the first transform
            float cv1[V];
            cot(gx, cv);
            cot(gx + 1, cv1);
            u_pair(j0);
#pragma unroll
            for (int j = 0; j < V; ++j) {
                wx[(2 * w + 1) * M + lane + kWave * j] = u[j].x * cv[j];
                wy0[j] = u[j].y * cv[j];
                cv[j] = cv1[j];
            }
            u_pair(j0 + 1);
#pragma unroll
            for (int j = 0; j < V; ++j) {
                wx[(2 * w + 2) * M + lane + kWave * j] = u[j].x * cv[j];
                // w_y of both columns in the line; the y-stencil transpose reads
                // the neighbouring rows (same wavefront, in order)
                line[lslot<M>(lane + kWave * j)] = make_float2(wy0[j], u[j].y * cv[j]);
            }
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int y = lane + kWave * j;
                const float2 wm = y - 1 >= 0 ? line[lslot<M>(y - 1)] : make_float2(0.f, 0.f);
                const float2 wp = y + 1 < M ? line[lslot<M>(y + 1)] : make_float2(0.f, 0.f);
                acc0[j] += 0.5f * (wm.x - wp.x);
                acc1[j] += 0.5f * (wm.y - wp.y);
            }
            if (w == P3A_WAVES - 1) {
                // halo columns jb - 1 and jb + CBK: u_x only
                float cr[V];
                const int cl = t.ox0 + jb - 1, cr0 = t.ox0 + jb + CBK;
                cot(cl, cv);
                cot(cr0, cr);
                if ((cl >= e && cl < N - e) || (cr0 >= e && cr0 < N - e)) {
                    ifft_pair<M, PL, float2, COH, true>(FP, line, col(px, jb - 1),
                                                        col(px, jb + CBK), u, lane);
                } else {
#pragma unroll
                    for (int q = 0; q < V; ++q) u[q] = make_float2(0.f, 0.f);
                }
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    wx[lane + kWave * j] = u[j].x * cv[j];
                    wx[(CBK + 1) * M + lane + kWave * j] = u[j].y * cr[j];
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int y = lane + kWave * j;
                acc0[j] += 0.5f * (wx[(2 * w) * M + y] - wx[(2 * w + 2) * M + y]);
                acc1[j] += 0.5f * (wx[(2 * w + 1) * M + y] - wx[(2 * w + 3) * M + y]);
            }
            __syncthreads();
        }
        if (j0 < cnt) {
            const bool second = j0 + 1 < cnt;
            float* o0 = a.out0 + ib + (long long)gx * N;
#pragma unroll
            for (int j = 0; j < V; ++j) {
                const int gy = t.p0y + lane + kWave * j;
                if (gy >= t.oy0 && gy < t.oy1) {
                    stu(o0, (unsigned)gy, acc0[j]);
                    if (second) stu(o0 + N, (unsigned)gy, acc1[j]);
                }
            }
        }
    }
    // lambda_g
    if (j0 < cnt) {
        const bool second = j0 + 1 < cnt;
        float2 l[V];
        ifft_pair<M, PL, SL, COH, true>(FP, line, colS(SL{}, planeL, j0),
                                        colS(SL{}, planeL, j0 + 1), l, lane);
        float* o1 = a.out1 + ib + (long long)gx * N;
#pragma unroll
        for (int j = 0; j < V; ++j) {
            const int gy = t.p0y + lane + kWave * j;
            if (gy >= t.oy0 && gy < t.oy1) {
                stu(o1, (unsigned)gy, l[j].x);
                if (second) stu(o1 + N, (unsigned)gy, l[j].y);
            }
        }
    }
}

template <int M, class PL, int NF, class SL>
__global__ __launch_bounds__(P3A_WAVES * kWave)
__attribute__((amdgpu_waves_per_eu(M <= 512 && NF == 1 && !IsMfmaPlan<PL>::value ? 4 : 1))) void k_p3_adj(EngineArgs a, int with_u, int planeL)
{
    __shared__ float2 lds[P3A_WAVES][line_slots(M)];
    __shared__ float wx[p3a_wx_floats<M>()];
    p3_adj_item<M, PL, NF, SL>(a, with_u, planeL, a.tiles[blockIdx.y], blockIdx.y, blockIdx.x,
                               lds, wx);
}

// lambda_f from caller-provided u_x, u_y [W x W] (adjoint/srk_adjoint.cuh:25-59).
__global__ void k_adjoint_f_from_u(const float* __restrict__ v, const float* __restrict__ ux,
                                   const float* __restrict__ uy, float* __restrict__ lam,
                                   int N, int W, int e, long long img_stride, long long w_stride)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int b = blockIdx.z;
    if (r >= N) return;
    const float* vb = v + b * w_stride;
    const float* uxb = ux + b * w_stride;
    const float* uyb = uy + b * w_stride;
    auto wx = [&](int rr, int cc) -> float {
        const int orr = rr - e, oc = cc - e;
        if (orr < 0 || orr >= W || oc < 0 || oc >= W) return 0.f;
        return uxb[(long long)oc * W + orr] * vb[(long long)oc * W + orr];
    };
    auto wy = [&](int rr, int cc) -> float {
        const int orr = rr - e, oc = cc - e;
        if (orr < 0 || orr >= W || oc < 0 || oc >= W) return 0.f;
        return uyb[(long long)oc * W + orr] * vb[(long long)oc * W + orr];
    };
    lam[b * img_stride + (long long)c * N + r] =
        0.5f * (wx(r, c - 1) - wx(r, c + 1)) + 0.5f * (wy(r - 1, c) - wy(r + 1, c));
}

// GT[d][k_y][k_x] = scale * G_d[k_x*(M/2+1) + k_y]  (32 x 32 LDS-tiled transpose)
__global__ void k_import_spectra(const float2* __restrict__ Gx, const float2* __restrict__ Gy,
                                 float2* __restrict__ GT, int M, float scale)
{
    __shared__ float2 tile[32][33];
    const int H = M / 2 + 1;
    const float2* src = blockIdx.z == 0 ? Gx : Gy;
    float2* dst = GT + (long long)blockIdx.z * H * M;
    const int kx0 = blockIdx.x * 32, ky0 = blockIdx.y * 32;
    for (int i = threadIdx.y; i < 32; i += blockDim.y) {
        const int kx = kx0 + i, ky = ky0 + threadIdx.x;
        if (kx < M && ky < H) tile[i][threadIdx.x] = src[(long long)kx * H + ky];
    }
    __syncthreads();
    for (int i = threadIdx.y; i < 32; i += blockDim.y) {
        const int ky = ky0 + i, kx = kx0 + threadIdx.x;
        if (kx < M && ky < H) {
            const float2 g = tile[threadIdx.x][i];
            dst[(long long)ky * M + kx] = make_float2(g.x * scale, g.y * scale);
        }
    }
}

// ---------------------------------------------------------------------------
// Fused pipeline: the three passes of Op::Forward or Op::Adjoint for batches
// of one-tile images (N <= M), in one persistent launch.  Each workgroup reads
// the hardware id of its XCD and serves that XCD's queue: the XCD owns images
// b = 8 j + xcd (j = 0, 1, ...), and its exchange planes live in a ring of K
// slots per XCD (image j in slot j mod K), so an image's planes are written
// and read through that XCD's L2 and the Infinity Cache instead of
// round-tripping HBM.
//
// Queue order: round r holds the stage-3 items (column inverse transforms)
// of image r - 2L, then the stage-2 items (row passes) of image r - L, then
// the stage-1 items (column transforms) of image r.  Stage 1 of image j waits
// for stage 2 of image j - K (its X1 slot), stage 2 of j for stage 1 of j and
// for stage 3 of j - K (its X2 slot), stage 3 of j for stage 2 of j.  With
// K > L every item depends only on items of earlier rounds, so items claimed
// in order can never wait on an item no running workgroup holds; with
// K >= L + 2 and L >= 2 the items waited on were claimed two or more rounds
// earlier, so in steady state they are complete and no round waits on the
// stragglers of the previous one.  A stage publishes its planes by waiting
// for its stores to complete (they are then in the XCD's L2), a barrier, and
// one relaxed agent-scope increment of the slot's counter; a consumer polls
// the counter (relaxed agent-scope loads) and reads the planes with coherent
// loads that miss its CU's L1.  Every wait is bounded by a deadline: past it
// the workgroup raises the abort flag and counts a timeout
// (srk::fused_sync_status), and all workgroups drain the queue without work.
//
// Every value that steers the loop (claimed item, wait results) is made
// wavefront-uniform (readfirstlane) and the loop has a single latch, and the
// claim and the polls run in wave-uniform branches.  When the item index was
// read from LDS as a per-lane value and the claim, the polls and the
// deadline check were done by thread 0 alone, the compiler built the loop as
// a per-lane loop in which thread 0 left to claim the next item while the
// other lanes re-ran the last one without end (the workgroup never reached a
// barrier it needed, and the deadline check was on thread 0's path only).
// ---------------------------------------------------------------------------
#ifndef FUSED_WPE
#define FUSED_WPE 4
#endif
constexpr int kSyncLine = 32;  // ints per 128-byte line

struct FusedSync {
    int abort[kSyncLine];
    int status[kSyncLine];                      // [0] timed-out waits
    int x[8][1 + 3 * kFusedSlotsMax][kSyncLine];  // per XCD: claim counter, c1[K], c2[K], c3[K]
};

struct FusedArgs {
    EngineArgs a;                // X1 / X2: rings of 8 K slots (slot = K xcd + j mod K)
    FusedSync* sync;
    int K, L;                    // ring slots per XCD, stage lag in rounds
    int B;                       // images
    int n1a, n1b;                // stage 1: column blocks of P1 (n1a), then of P1A (n1b)
    int bx1b;                    // first column block of the P1A part
    int n2;                      // stage 2: bands per row pass (NBF + 1)
    int n3;                      // stage 3: column blocks of P3
    unsigned long long budget;   // wall-clock ticks (100 MHz) a workgroup may run
};

__device__ __forceinline__ int xcd_id()
{
    return __builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 7;  // HW_REG_XCC_ID[3:0]
}

__device__ __forceinline__ int sync_load(int* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ long long uni(long long v)
{
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)v >> 32));
    return (long long)(((unsigned long long)hi << 32) | lo);
}

// index of the calling wavefront, uniform in the compiler's view
__device__ __forceinline__ int wave_index() { return uni((int)(threadIdx.x >> 6)); }

// Wait until *ctr >= target; false after the deadline or an abort.  Wavefront 0
// polls with uniform control (every lane loads the counter, the value is made
// uniform), so the poll loop is a scalar loop; the other wavefronts wait at
// the barrier and learn the result through LDS.
__device__ bool fused_wait(FusedSync* sy, int* ctr, int target, unsigned long long t_end,
                           int* flag)
{
    if (wave_index() == 0) {
        int ok = 1;
        for (;;) {
            if (uni(sync_load(ctr)) >= target) break;
            if (uni(sync_load(&sy->abort[0])) != 0) {
                ok = 0;
                break;
            }
            if (wall_clock64() > t_end) {
                if (threadIdx.x == 0) {
                    __hip_atomic_fetch_add(&sy->status[0], 1, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&sy->abort[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        *flag = ok;  // every lane of wavefront 0 writes the same value
    }
    __syncthreads();
    const int ok = uni(*flag);
    __syncthreads();  // every wavefront has read the flag before it is rewritten
    return ok != 0;
}

// publish this block's stores: complete them (into the L2), then count the item
__device__ __forceinline__ void fused_signal(int* ctr)
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int M>
constexpr int fused_lds_bytes()
{
    constexpr int lines = 4 * line_slots(M) * 8;                        // P1, P1A, P3 fwd
    constexpr int p2 = kFusedRY * p2_pitch<M, kFusedRY>() * 8;           // row passes
    constexpr int p3a = P3A_WAVES * line_slots(M) * 8 + p3a_wx_floats<M>() * 4;
    return lines > p2 ? (lines > p3a ? lines : p3a) : (p2 > p3a ? p2 : p3a);
}

// clears the counters and writes the arguments the persistent launch reads
// (kernel arguments would stay live in registers across every item type)
__global__ void k_fused_setup(FusedArgs fa, FusedArgs* dst)
{
    int* p = reinterpret_cast<int*>(fa.sync);
    for (int i = threadIdx.x; i < (int)(sizeof(FusedSync) / sizeof(int)); i += blockDim.x) p[i] = 0;
    if (threadIdx.x == 0) *dst = fa;
}

// Values the persistent loop uses are re-read from LDS for every item and
// made wavefront-uniform (readfirstlane), so that nothing derived from the
// arguments is loop-invariant: otherwise the compiler hoists every item kind's
// address arithmetic and twiddle loads out of the loop and keeps them all live
// at once.  Pointers come back as global-address-space pointers (they are
// device allocations), so the item code keeps its scalar-base global loads.
template <class T>
__device__ __forceinline__ T* uni_global(T* p)
{
    typedef __attribute__((address_space(1))) T gT;
    return (T*)(gT*)(unsigned long long)uni((long long)(unsigned long long)p);
}

__device__ __forceinline__ EngineArgs args_from_lds(const EngineArgs& s)
{
    EngineArgs a;
    a.g0 = uni_global(s.g0);
    a.g1 = uni_global(s.g1);
    a.f0 = uni_global(s.f0);
    a.f1 = uni_global(s.f1);
    a.v0 = uni_global(s.v0);
    a.v1 = uni_global(s.v1);
    a.out0 = uni_global(s.out0);
    a.out1 = uni_global(s.out1);
    a.ux_out = uni_global(s.ux_out);
    a.uy_out = uni_global(s.uy_out);
    a.img_stride = uni(s.img_stride);
    a.w_stride = uni(s.w_stride);
    a.N = uni(s.N);
    a.e = uni(s.e);
    a.W = uni(s.W);
    a.GT = uni_global(s.GT);
    a.tw = uni_global(s.tw);
    a.zcol = uni_global(s.zcol);
    a.tiles = uni_global(s.tiles);
    a.X1 = uni_global(s.X1);
    a.X2 = uni_global(s.X2);
    a.np1 = uni(s.np1);
    a.np2 = uni(s.np2);
    a.S2 = uni(s.S2);
    a.halo2 = uni(s.halo2);
    a.p2_rows = uni(s.p2_rows);
    a.ry_log2 = uni(s.ry_log2);
    return a;
}

// The loop's control values (claimed item, wait results) are uniform
// (readfirstlane) and the loop has a single latch: with per-lane control flow
// (an item index read from LDS, `continue` paths, a claim and a signal done by
// thread 0 only) the compiler may split the loop by lane, leaving all lanes
// but thread 0 in an inner loop that re-runs the last item and never reaches
// the claim.
template <int M, int OP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FUSED_WPE)))
void k_fused(const FusedArgs* fap)
{
    static_assert(P3A_WAVES == 4, "the fused workgroup is four wavefronts");
    using PL = FftPlan<M>;
    using S = float2;
    constexpr bool ADJ = OP == (int)Op::Adjoint;
    __shared__ __align__(16) unsigned char smem[fused_lds_bytes<M>()];
    __shared__ FusedArgs s_fa;
    __shared__ int s_item, s_flag;
    float2(*lines)[line_slots(M)] = reinterpret_cast<float2(*)[line_slots(M)]>(smem);
    float2* band_lines = reinterpret_cast<float2*>(smem);
    float* wx = reinterpret_cast<float*>(smem + P3A_WAVES * line_slots(M) * 8);

    {
        const int* src = reinterpret_cast<const int*>(fap);
        int* dst = reinterpret_cast<int*>(&s_fa);
        for (int k = threadIdx.x; k < (int)(sizeof(FusedArgs) / sizeof(int)); k += blockDim.x)
            dst[k] = src[k];
    }
    __syncthreads();
    FusedSync* sy = uni_global(s_fa.sync);
    const int x = xcd_id();
    int* q = sy->x[x][0];
    int* c1 = sy->x[x][1];   // c1 + kSyncLine * s: slot s
    int* c2 = sy->x[x][1 + kFusedSlotsMax];
    int* c3 = sy->x[x][1 + 2 * kFusedSlotsMax];
    const int K = uni(s_fa.K), L = uni(s_fa.L);
    const int B = uni(s_fa.B);
    const int nx = B > x ? (B - x + 7) / 8 : 0;  // images of this XCD
    const int n1a = uni(s_fa.n1a), n2a = uni(s_fa.n2), bx1b = uni(s_fa.bx1b);
    const int n1 = n1a + uni(s_fa.n1b), n2 = (ADJ ? 2 : 1) * n2a, n3 = uni(s_fa.n3);
    const int ipr = n1 + n2 + n3;
    const int total = (nx + 2 * L) * ipr;
    const unsigned long long t_end = wall_clock64() + (unsigned long long)uni((long long)s_fa.budget);
    for (int i = 0; nx > 0;) {
        if (wave_index() == 0) {
            int v = 0;
            if (threadIdx.x == 0)
                v = __hip_atomic_fetch_add(q, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            v = uni(v);
            if (uni(sync_load(&sy->abort[0])) != 0 || wall_clock64() > t_end) v = -1;
            s_item = v;
        }
        __syncthreads();
        i = uni(s_item);
        __syncthreads();  // every wavefront has read s_item before it is rewritten
        if (i < 0 || i >= total) break;
        const int r = i / ipr, k = i - r * ipr;
        const int stage = k < n3 ? 3 : k < n3 + n2 ? 2 : 1;
        const int j = r - (stage == 3 ? 2 * L : stage == 2 ? L : 0);
        if (j >= 0 && j < nx) {  // else pipeline fill / drain: nothing to do
            const int idx = stage == 3 ? k : stage == 2 ? k - n3 : k - n3 - n2;
            const int g = j / K, s = j - g * K;
            const int slot = K * x + s;
            const EngineArgs a = args_from_lds(s_fa.a);
            const int b = 8 * j + x;
            const Tile t = ADJ ? Tile{b, 0, 0, 0, a.N, 0, a.N, 0}
                               : Tile{b, 0, 0, a.e, a.N - a.e, a.e, a.N - a.e, 0};
            if (stage == 1) {
                if (fused_wait(sy, c2 + kSyncLine * s, n2 * g, t_end, &s_flag)) {
                    if (idx < n1a)
                        p1_cols_item<M, PL, 1, S, false>(a, t, slot, idx, lines);
                    else
                        p1_adj_src_item<M, PL, 1, S, false>(a, 1, t, slot, bx1b + idx - n1a, lines);
                    fused_signal(c1 + kSyncLine * s);
                }
            } else if (stage == 2) {
                if (fused_wait(sy, c1 + kSyncLine * s, n1 * (g + 1), t_end, &s_flag) &&
                    fused_wait(sy, c3 + kSyncLine * s, n3 * g, t_end, &s_flag)) {
                    if (idx < n2a)
                        p2_conv_band_fused<M, PL, S>(a, t, slot, idx, band_lines);
                    else
                        p2_corr_band_fused<M, PL, S>(a, t, slot, idx - n2a, 1, 2, band_lines);
                    fused_signal(c2 + kSyncLine * s);
                }
            } else {
                if (fused_wait(sy, c2 + kSyncLine * s, n2 * (g + 1), t_end, &s_flag)) {
                    if constexpr (ADJ)
                        p3_adj_item<M, PL, 1, S, true>(a, 1, 2, t, slot, idx, lines, wx);
                    else
                        p3_fwd_item<M, PL, 1, S, true>(a, t, slot, idx, lines);
                    fused_signal(c3 + kSyncLine * s);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <int M>
struct P2Cfg;
template <>
struct P2Cfg<256> { static constexpr int RY = 8, NW = 8; };
template <>
struct P2Cfg<512> { static constexpr int RY = 8, NW = 8; };
template <>
struct P2Cfg<1024> { static constexpr int RY = 8, NW = 8; };
template <>
struct P2Cfg<2048> { static constexpr int RY = 4, NW = 4; };

// Wavefronts per P2 block: one LDS line per wavefront, except that the MFMA
// plan holds more per-lane constants and runs two lines per wavefront at
// M >= 512 so that a lane may use 256 registers.
template <int M, int RY, class PL>
struct P2Waves {
    static constexpr int NW = (IsMfmaPlan<PL>::value && M >= 512 && RY >= 8) ? RY / 2 : RY;
};

template <int M, int RY, class PL, int PASS>
constexpr size_t p2_lds()
{
    return (p2_one_line<M, PL, PASS>() ? 1ull : 2ull) * RY * p2_pitch<M, RY>() * sizeof(float2);
}

// The dynamic-LDS limits are per-kernel attributes: set once per process
// (a function-local static is initialised exactly once, also under
// concurrent first calls).
template <int M, int RY, class PL, class SX>
hipError_t prepare_kernels()
{
    static const hipError_t st = [] {
        constexpr int NW = P2Waves<M, RY, PL>::NW;
        constexpr int lc = (int)p2_lds<M, RY, PL, kP2ConvPass>();
        constexpr int lr = (int)p2_lds<M, RY, PL, kP2CorrPass>();
        const struct { const void* fn; int lds; } fns[] = {
            {reinterpret_cast<const void*>(&k_p2_conv<M, RY, NW, PL, float2>), lc},
            {reinterpret_cast<const void*>(&k_p2_conv<M, RY, NW, PL, SX>), lc},
            {reinterpret_cast<const void*>(&k_p2_corr<M, RY, NW, PL, SX>), lr}};
        hipError_t e = hipSuccess;
        for (const auto& f : fns)
            if (e == hipSuccess)
                e = hipFuncSetAttribute(f.fn, hipFuncAttributeMaxDynamicSharedMemorySize, f.lds);
        return e;
    }();
    return st;
}

// SX is the element format of the forward and tangent planes and of the
// correlation planes of the adjoints; the adjoints' convolution planes (the
// recomputed u, which lambda_f multiplies by the cotangent) stay FP32.
template <int M, class PL, int RY, class SX>
int run_impl(Op op, const EngineArgs& a0, int ntiles, int chunk, hipStream_t s0, LaunchTimer* tm,
             const Fork* fk)
{
    static_assert(RY == 16 || RY == 8 || RY == 4, "band heights");
    constexpr int NW = P2Waves<M, RY, PL>::NW;
    hipError_t st = prepare_kernels<M, RY, PL, SX>();
    if (st != hipSuccess) return (int)st;
    const size_t lds2c = p2_lds<M, RY, PL, kP2ConvPass>();
    const size_t lds2r = p2_lds<M, RY, PL, kP2CorrPass>();
    const int cols_out = a0.S2 - 2 * a0.halo2;  // max output columns per tile
    bool aux_used = false;
    for (int t0 = 0, ci = 0; t0 < ntiles; t0 += chunk, ++ci) {
        const int c = std::min(chunk, ntiles - t0);
        EngineArgs a = a0;
        a.tiles = a0.tiles + t0;
        // lane of this chunk: odd chunks run on the second lane when there is one
        const bool lane1 = fk && (ci & 1);
        const hipStream_t s = lane1 ? fk->aux : s0;
        if (lane1) {
            a.X1 = fk->X1b;
            a.X2 = fk->X2b;
            if (!aux_used && st == hipSuccess) {
                st = hipEventRecord(fk->fork, s0);
                if (st == hipSuccess) st = hipStreamWaitEvent(fk->aux, fk->fork, 0);
                aux_used = true;
            }
        }
        a.ry_log2 = RY == 16 ? 4 : RY == 8 ? 3 : 2;
        const int nsrc = a.src_b1 > a.src_b0 ? a.src_b1 - a.src_b0 : M / 4;
        const dim3 g1(M / 8, c), g1a(nsrc, c);
        const dim3 g3f((cols_out + 3) / 4, c), g3a((cols_out + 2 * P3A_WAVES - 1) / (2 * P3A_WAVES), c);
        const dim3 b256(256), b2(NW * kWave), b3a(P3A_WAVES * kWave);
        // launch one pass on stream ls, bracketed by timing events when a timer is given
        auto run = [&](int kind, hipStream_t ls, auto&& launch) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (tm && st == hipSuccess) st = tm->record(ls, &e0);
            launch();
            if (tm && st == hipSuccess) {
                st = tm->record(ls, &e1);
                tm->recs.push_back({kind, c, e0, e1});
            }
        };
        // x16: the convolution planes of this pass use the format SX (else FP32)
        auto p1 = [&](int nf, bool x16, hipStream_t ls) {
            run(kP1Cols, ls, [&] {
                if (nf == 1 && x16)
                    k_p1_cols<M, PL, 1, SX><<<g1, b256, 0, ls>>>(a);
                else if (nf == 1)
                    k_p1_cols<M, PL, 1, float2><<<g1, b256, 0, ls>>>(a);
                else if (x16)
                    k_p1_cols<M, PL, 2, SX><<<g1, b256, 0, ls>>>(a);
                else
                    k_p1_cols<M, PL, 2, float2><<<g1, b256, 0, ls>>>(a);
            });
        };
        auto p1a = [&](int nf, int plane0, hipStream_t ls) {
            run(kP1AdjSrc, ls, [&] {
                if (nf == 1)
                    k_p1_adj_src<M, PL, 1, SX><<<g1a, b256, 0, ls>>>(a, plane0);
                else
                    k_p1_adj_src<M, PL, 2, SX><<<g1a, b256, 0, ls>>>(a, plane0);
            });
        };
        auto p2c = [&](int nf, bool x16, hipStream_t ls) {
            const dim3 g2x(P2Items<M, RY, p2c_items<M, PL, SX>()>::blocks(c, nf));
            const dim3 g2f(P2Items<M, RY, p2c_items<M, PL, float2>()>::blocks(c, nf));
            run(kP2Conv, ls, [&] {
                if (x16)
                    k_p2_conv<M, RY, NW, PL, SX><<<g2x, b2, lds2c, ls>>>(a, nf, c);
                else
                    k_p2_conv<M, RY, NW, PL, float2><<<g2f, b2, lds2c, ls>>>(a, nf, c);
            });
        };
        auto p2r = [&](int in, int out, hipStream_t ls) {
            const dim3 g2(P2Items<M, RY, 1>::blocks(c, 1));
            run(kP2Corr, ls,
                [&] { k_p2_corr<M, RY, NW, PL, SX><<<g2, b2, lds2r, ls>>>(a, in, out, c); });
        };
        auto p3f = [&](int nf) {
            run(kP3Fwd, s, [&] {
                if (nf == 1)
                    k_p3_fwd<M, PL, 1, SX><<<g3f, b256, 0, s>>>(a);
                else
                    k_p3_fwd<M, PL, 2, SX><<<g3f, b256, 0, s>>>(a);
            });
        };
        auto p3a = [&](int nf, int with_u, int planeL) {
            run(kP3Adj, s, [&] {
                if (nf == 1)
                    k_p3_adj<M, PL, 1, SX><<<g3a, b3a, 0, s>>>(a, with_u, planeL);
                else
                    k_p3_adj<M, PL, 2, SX><<<g3a, b3a, 0, s>>>(a, with_u, planeL);
            });
        };
        switch (op) {
        case Op::Forward:
            // u requested as an output (for a later adjoint) is computed in FP32
            p1(1, a.ux_out == nullptr, s);
            p2c(1, a.ux_out == nullptr, s);
            if (a.ux_out)
                run(kP3Fwd, s, [&] { k_p3_fwd<M, PL, 1, float2><<<g3f, b256, 0, s>>>(a); });
            else
                p3f(1);
            break;
        case Op::Tangent:
            p1(2, true, s);
            p2c(2, true, s);
            p3f(2);
            break;
        case Op::Adjoint:
            p1(1, false, s);
            p1a(1, 1, s);
            p2c(1, false, s);
            p2r(1, 2, s);
            p3a(1, 1, 2);
            break;
        case Op::AdjointG:
            p1a(1, 0, s);
            p2r(0, 0, s);
            p3a(1, 0, 0);
            break;
        case Op::TangentAdjoint:
            p1(2, false, s);
            p1a(2, 2, s);
            p2c(2, false, s);
            p2r(2, 4, s);
            p3a(2, 1, 4);
            break;
        }
        if (st == hipSuccess) st = hipGetLastError();
        if (st != hipSuccess) break;
    }
    // the caller's stream always ends after the second lane's work, on the
    // error path too (a capture must not be left with an unjoined stream)
    if (aux_used) {
        const hipError_t j1 = hipEventRecord(fk->join, fk->aux);
        const hipError_t j2 = j1 == hipSuccess ? hipStreamWaitEvent(s0, fk->join, 0) : j1;
        if (st == hipSuccess) st = j2;
    }
    return (int)st;
}

// P2 band height: the tuned value if it is one of the compiled variants,
// otherwise the default for M.
template <int M, class PL, class SX = float2>
int run_rows(Op op, const EngineArgs& a, int ntiles, int chunk, int rows, hipStream_t s,
             LaunchTimer* tm, const Fork* fk)
{
    if constexpr (M <= 512) {
        if (rows == 16) return run_impl<M, PL, 16, SX>(op, a, ntiles, chunk, s, tm, fk);
        if (rows == 4) return run_impl<M, PL, 4, SX>(op, a, ntiles, chunk, s, tm, fk);
    }
    return run_impl<M, PL, P2Cfg<M>::RY, SX>(op, a, ntiles, chunk, s, tm, fk);
}

}  // namespace

int engine_run(Op op, int M, int prec, const EngineArgs& a, int ntiles, int chunk, hipStream_t s,
               LaunchTimer* tm, const Fork* fk)
{
    if (chunk < 1 || ntiles < 0) return (int)hipErrorInvalidValue;
    if (ntiles == 0) return 0;
    const int rows = a.p2_rows;
    if (prec == 0) {
        switch (M) {
        case 256: return run_rows<256, FftPlan<256>>(op, a, ntiles, chunk, rows, s, tm, fk);
        case 512: return run_rows<512, FftPlan<512>>(op, a, ntiles, chunk, rows, s, tm, fk);
        case 1024: return run_rows<1024, FftPlan<1024>>(op, a, ntiles, chunk, rows, s, tm, fk);
        case 2048: return run_rows<2048, FftPlan<2048>>(op, a, ntiles, chunk, rows, s, tm, fk);
        default: return (int)hipErrorInvalidValue;
        }
    }
    if (prec == 1) {
        switch (M) {
        case 256: return run_rows<256, FftPlanMfma<256>>(op, a, ntiles, chunk, rows, s, tm, fk);
        case 512: return run_rows<512, FftPlanMfma<512>>(op, a, ntiles, chunk, rows, s, tm, fk);
        case 1024: return run_rows<1024, FftPlanMfma<1024>>(op, a, ntiles, chunk, rows, s, tm, fk);
        default: return (int)hipErrorInvalidValue;
        }
    }
    if (prec == 2) {
        using P = Bf16c;
        switch (M) {
        case 256: return run_rows<256, FftPlan<256>, P>(op, a, ntiles, chunk, rows, s, tm, fk);
        case 512: return run_rows<512, FftPlan<512>, P>(op, a, ntiles, chunk, rows, s, tm, fk);
        case 1024: return run_rows<1024, FftPlan<1024>, P>(op, a, ntiles, chunk, rows, s, tm, fk);
        case 2048: return run_rows<2048, FftPlan<2048>, P>(op, a, ntiles, chunk, rows, s, tm, fk);
        default: return (int)hipErrorInvalidValue;
        }
    }
    return (int)hipErrorInvalidValue;
}

size_t fused_sync_bytes() { return sizeof(FusedSync) + sizeof(FusedArgs); }

bool fused_supported(Op op, int M, int N, int prec)
{
    return prec == 0 && (op == Op::Forward || op == Op::Adjoint) && (M == 256 || M == 512) &&
           N <= M;
}

namespace {

// persistent grid: as many workgroups as are resident at once, a multiple of 8
template <int M, int OP>
int fused_grid()
{
    static const int g = [] {
        int dev = 0, cus = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per, reinterpret_cast<const void*>(&k_fused<M, OP>), 256, 0) != hipSuccess)
            return 0;
        const int n = (cus * per) & ~7;
        return n > 0 ? n : 8;
    }();
    return g;
}

template <int M, int OP>
int launch_fused(const FusedArgs* fap, hipStream_t s)
{
    const int g = fused_grid<M, OP>();
    if (g <= 0) return (int)hipErrorInvalidValue;
    k_fused<M, OP><<<g, 256, 0, s>>>(fap);
    return (int)hipGetLastError();
}

}  // namespace

int engine_run_fused(Op op, int M, const EngineArgs& a, int B, void* sync, int slots, int lag,
                     hipStream_t s, LaunchTimer* tm)
{
    if (!fused_supported(op, M, a.N, 0) || B < 1 || !sync || lag < 1 || slots <= lag ||
        slots > kFusedSlotsMax)
        return (int)hipErrorInvalidValue;
    FusedArgs fa{};
    fa.a = a;
    fa.a.ry_log2 = 3;  // kFusedRY
    fa.sync = static_cast<FusedSync*>(sync);
    fa.K = slots;
    fa.L = lag;
    fa.B = B;
    const bool adj = op == Op::Adjoint;
    const int N = a.N, e = a.e, W = a.W;
    fa.n1a = M / 8;
    if (adj) {
        const int lo = e >= 3 ? e / 4 : 0;              // first block with 4 bx + 3 >= e
        const int hi = std::min(M / 4, (N - e + 3) / 4);  // blocks with 4 bx < N - e
        fa.bx1b = lo;
        fa.n1b = std::max(0, hi - lo);
    }
    fa.n2 = M / (2 * kFusedRY) + 1;
    fa.n3 = adj ? (N + 2 * P3A_WAVES - 1) / (2 * P3A_WAVES) : (W + 3) / 4;
    // wall-clock budget of a workgroup (100 MHz ticks): 2 s plus 0.1 ms per
    // image, far above the few microseconds an image takes
    fa.budget = 200000000ull + 10000ull * (unsigned long long)B;
    FusedArgs* fap = reinterpret_cast<FusedArgs*>(static_cast<char*>(sync) + sizeof(FusedSync));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t st = hipSuccess;
    if (tm) st = tm->record(s, &e0);
    k_fused_setup<<<1, 256, 0, s>>>(fa, fap);
    if (st == hipSuccess) st = hipGetLastError();
    if (st == hipSuccess) {
        int r;
        if (M == 256)
            r = adj ? launch_fused<256, (int)Op::Adjoint>(fap, s) : launch_fused<256, (int)Op::Forward>(fap, s);
        else
            r = adj ? launch_fused<512, (int)Op::Adjoint>(fap, s) : launch_fused<512, (int)Op::Forward>(fap, s);
        st = (hipError_t)r;
    }
    if (tm && st == hipSuccess) {
        st = tm->record(s, &e1);
        tm->recs.push_back({kFused, B, e0, e1});
    }
    return (int)st;
}

int fused_sync_status(const void* sync, int* timeouts)
{
    int h[2] = {0, 0};
    const hipError_t st = hipMemcpy(h, static_cast<const FusedSync*>(sync)->status, sizeof(h),
                                    hipMemcpyDeviceToHost);
    if (timeouts) *timeouts = h[0];
    return (int)st;
}

int engine_import_spectra(int M, const float2* Gx_ref, const float2* Gy_ref, float scale,
                          float2* GT, hipStream_t s)
{
    const int H = M / 2 + 1;
    dim3 grid((M + 31) / 32, (H + 31) / 32, 2), block(32, 8);
    k_import_spectra<<<grid, block, 0, s>>>(Gx_ref, Gy_ref, GT, M, scale);
    return (int)hipGetLastError();
}

int engine_adjoint_f_from_u(const float* v, const float* ux, const float* uy, float* lam_f,
                            int N, int W, int e, int B, long long img_stride,
                            long long w_stride, hipStream_t s)
{
    dim3 block(256), grid((N + 255) / 256, N, B);
    k_adjoint_f_from_u<<<grid, block, 0, s>>>(v, ux, uy, lam_f, N, W, e, img_stride, w_stride);
    return (int)hipGetLastError();
}

}  // namespace srk
