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

namespace srk {

namespace {

template <class PL>
struct IsMfmaPlan { static constexpr bool value = false; };
template <int M>
struct IsMfmaPlan<FftPlanMfma<M>> { static constexpr bool value = true; };

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
template <class S>
__device__ __forceinline__ void stx(S* base, unsigned idx, float2 v) { stu(base, idx, pack<S>(v)); }
// plane `index` of an exchange buffer whose planes hold `elems` FP32 elements
template <class S>
__device__ __forceinline__ S* plane_at(float2* X, long long index, long long elems)
{
    return reinterpret_cast<S*>(X + index * elems);
}

// Two Hermitian half spectra P, Q (X2 columns, or the zero column) as loaded
// for the lane's transform inputs k = lane + 64 i (raw, so that a load can be
// issued one transform ahead), and their packing Z = P + iQ.  The DC and
// Nyquist bins are taken as real, as a real inverse transform does.
template <int M, class S>
struct PairRaw {
    static constexpr int V = M / kWave;
    S p[V], q[V];

    __device__ __forceinline__ void load(const S* P, const S* Q, int lane)
    {
#pragma unroll
        for (int i = 0; i < V; ++i) {
            const int k = lane + kWave * i;
            const unsigned kk = k <= M / 2 ? k : M - k;
            p[i] = ldu(P, kk);
            q[i] = ldu(Q, kk);
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

// Inverse transform of P + iQ; afterwards out[i] = p(y) + i q(y) at tile row
// y = lane + 64 i.
template <int M, class PL, class S>
__device__ __forceinline__ void ifft_pair(const PL& FP, float2* line, const S* P, const S* Q,
                                          float2 (&out)[M / kWave], int lane)
{
    PairRaw<M, S> r;
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
template <int M, class PL, int NF, class S>
__global__ __launch_bounds__(256) void k_p1_cols(EngineArgs a)
{
    constexpr int nf = NF;
    constexpr int V = M / kWave;
    constexpr int H = M / 2 + 1;
    constexpr int HP = spec_pitch(M);
    constexpr int KV = (H + kWave - 1) / kWave;  // half-spectrum values per lane
    __shared__ float2 lds[4][line_slots(M)];
    const int lane = lane_id(), w = wave_id();
    PL FP;
    FP.load(a.tw, lane);
    const int slot = blockIdx.y;
    const Tile t = a.tiles[slot];
    const int N = a.N;
    const int x0 = blockIdx.x * 8;
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
            if (va && vb && t.p0y >= 0 && t.p0y + M <= N) {
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    const int row = t.p0y + lane + kWave * i;
                    v[i] = make_float2(ldu(colA, row), ldu(colB, row));
                }
            } else {
#pragma unroll
                for (int i = 0; i < V; ++i) {
                    const int row = t.p0y + lane + kWave * i;
                    const bool rv = row >= 0 && row < N;
                    const int rc = min(max(row, 0), N - 1);
                    const float pa = ldu(colA, rc), pb = ldu(colB, rc);
                    v[i] = make_float2((va && rv) ? pa : 0.f, (vb && rv) ? pb : 0.f);
                }
            }
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
            stx(plane, ((b * M + x0) << a.ry_log2) + (e & (8 * RY - 1)),
                lds[c >> 1][(c & 1) * H + b * RY + r]);
        }
        if (threadIdx.x < 8)  // the Nyquist row k = M/2
            stx(plane, x1_index<M>(M / 2, x0 + threadIdx.x, a.ry_log2),
                lds[threadIdx.x >> 1][(threadIdx.x & 1) * H + M / 2]);
    }
}

// ---------------------------------------------------------------------------
// P1A: adjoint source a_d = v0 * d_d f0 (+ v1 * d_d f1) on the valid region,
// transformed along y as a_x + i a_y; one wavefront per column, 4 per block,
// band stores staged in the wavefronts' own LDS lines as in P1.  Inside the
// valid region every stencil neighbour lies inside the image (e >= 1), so the
// loads need no bounds checks: rows outside it load at a clamped row and are
// masked.
// ---------------------------------------------------------------------------
template <int M, class PL, int NF, class S>
__global__ __launch_bounds__(256) void k_p1_adj_src(EngineArgs a, int plane0)
{
    constexpr int nf = NF;
    constexpr int V = M / kWave;
    constexpr int HP = spec_pitch(M);
    __shared__ float2 lds[4][line_slots(M)];
    const int lane = lane_id(), w = wave_id();
    const int slot = blockIdx.y;
    const Tile t = a.tiles[slot];
    const int N = a.N, e = a.e, W = a.W;
    // a block whose four columns all lie off the valid region has only zero
    // sources; P2 corr does not read those columns
    const int gb = t.p0x + blockIdx.x * 4;
    if (gb + 3 < e || gb >= N - e) return;
    const int x = blockIdx.x * 4 + w;
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
    const int x0 = blockIdx.x * 4;
    const int RY = 1 << a.ry_log2;
#pragma unroll
    for (int pl = 0; pl < 2; ++pl) {
        S* plane = plane_at<S>(a.X1, (long long)slot * a.np1 + plane0 + pl, (long long)M * HP);
        for (int e2 = threadIdx.x; e2 < 2 * M; e2 += 256) {
            const int b = e2 >> (a.ry_log2 + 2), c = (e2 >> a.ry_log2) & 3, r = e2 & (RY - 1);
            stx(plane, ((b * M + x0) << a.ry_log2) + (e2 & (4 * RY - 1)),
                lds[c][pl * H + b * RY + r]);
        }
        if (threadIdx.x < 4)  // the Nyquist row k = M/2
            stx(plane, x1_index<M>(M / 2, x0 + threadIdx.x, a.ry_log2),
                lds[threadIdx.x][pl * H + M / 2]);
    }
}

// ---------------------------------------------------------------------------
// P2 work items.  An item is one band (RY consecutive k_y rows) of one tile
// (and, for the conv pass, one field).  Full bands come first in launch
// order: block L runs on XCD L mod 8 (workgroups are dispatched to the eight
// XCDs round robin) and only takes bands b with b mod 8 = L mod 8, so every
// XCD reads only its own eighth of the spectra, which stays resident in its
// L2.  Each such block takes up to ITEMS consecutive items of its XCD's
// list (the conv pass prefetches the next item's band into registers while it
// transforms the current one).  The Nyquist rows (k_y = M/2, a band of one
// row) follow, one item per block.
// ---------------------------------------------------------------------------
template <int M, int RY, int ITEMS>
struct P2Items {
    static constexpr int NBF = M / (2 * RY);  // full bands per tile
    static constexpr int BPX = NBF / 8;       // full bands per tile and XCD
    static_assert(NBF % 8 == 0, "full bands must split evenly over the eight XCDs");
    static constexpr int full_blocks(int tiles, int nf)
    {
        return 8 * ((BPX * tiles * nf + ITEMS - 1) / ITEMS);
    }
    static constexpr int blocks(int tiles, int nf) { return full_blocks(tiles, nf) + tiles * nf; }

    int xcd, first, count, nf;
    bool nyquist;
    __device__ P2Items(int tiles, int nf_) : nf(nf_)
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
    // item k of this block -> tile slot, band, field
    __device__ void get(int k, int& slot, int& band, int& fi) const
    {
        const int i = first + k;
        fi = i % nf;
        const int j = i / nf;
        if (nyquist) {
            slot = j;
            band = NBF;
        } else {
            slot = j / BPX;
            band = xcd + 8 * (j % BPX);
        }
    }
    __device__ int rows(int band) const { return band < NBF ? RY : 1; }
};

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

    // columns outside [xlo, xhi) (tile-local) are taken as zero and not read
    __device__ __forceinline__ void fetch(const S* __restrict__ band, int nrows, int xlo = 0,
                                          int xhi = M)
    {
        constexpr int XS = T / RY;
        const bool ok = (int)(threadIdx.x % RY) < nrows;
        const int x0 = threadIdx.x / RY;
        const S* src = band + threadIdx.x;
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int x = x0 + i * XS;
            v[i] = pack<S>(make_float2(0.f, 0.f));
            if (ok && x >= xlo && x < xhi) v[i] = ldu(src, i * T);
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

// LDS lines -> X2 plane ([x][k_y], pitch HP): X2 column xi holds tile-local
// column lo + xi; columns xi in [xs, xe) are stored, those outside [0, M) as
// zero.
template <int M, int RY, int NW, class S>
__device__ __forceinline__ void store_band(S* __restrict__ plane, const float2* L, int k0,
                                           int nrows, int lo, int xs, int xe)
{
    constexpr int T = NW * kWave, XS = T / RY, HP = spec_pitch(M), LP = p2_pitch<M, RY>();
    const int r = threadIdx.x % RY, xi0 = xs + threadIdx.x / RY;
    if (r >= nrows) return;
    const float2* src = L + r * LP;
    S* d = plane + (long long)xi0 * HP + k0 + r;
    for (int xi = xi0; xi < xe; xi += XS, d += (long long)XS * HP) {
        const int x = lo + xi;
        float2 v = make_float2(0.f, 0.f);
        if (x >= 0 && x < M) v = src[lslot<M>(x)];
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
enum P2Pass { kP2ConvPass = 0, kP2CorrPass = 1 };
template <int M, class PL, int PASS>
constexpr bool p2_one_line()
{
    return !IsMfmaPlan<PL>::value && M >= 1024;
}
template <int M, class PL>
constexpr int p2c_items() { return p2_one_line<M, PL, kP2ConvPass>() ? 1 : 2; }

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
constexpr int p2_waves_per_eu() { return p2_one_line<M, PL, PASS>() && M == 1024 ? 4 : 1; }

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
    extern __shared__ float2 sm[];
    float2* A = sm;
    float2* Bs = ONE ? sm : sm + RY * LP;
    const int lane = lane_id(), w = wave_id();
    PL FP;
    FP.load(a.tw, lane);
    const P2Items<M, RY, p2c_items<M, PL>()> items(tiles, nf);
    auto band_src = [&](int slot, int band, int fi) {
        return plane_at<const S>(a.X1, (long long)slot * a.np1 + fi, (long long)M * HP) +
               (long long)band * RY * M;
    };
    BandRegs<M, RY, NW, S> pf;
    int slot, band, fi;
    items.get(0, slot, band, fi);
    pf.fetch(band_src(slot, band, fi), items.rows(band));
    for (int k = 0; k < items.count; ++k) {
        items.get(k, slot, band, fi);
        const int k0 = band * RY;
        const int nrows = items.rows(band);
        if (k) __syncthreads();  // the previous item's stores have read A and Bs
        pf.put(A);
        __syncthreads();
        if (k + 1 < items.count) {
            int s2, b2, f2;
            items.get(k + 1, s2, b2, f2);
            pf.fetch(band_src(s2, b2, f2), items.rows(b2));
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
            float2 gxv[V], gyv[V];
#pragma unroll
            for (int i = 0; i < V; ++i) {
                gxv[i] = ldu(gxr, kWave * i);
                gyv[i] = ldu(gyr, kWave * i);
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
        items.get(k, slot, band, fi);
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
// P3 forward / tangent: inverse y-FFT of (u_x + i u_y) per output column,
// gradient stencil and combine, valid-region store.
// ---------------------------------------------------------------------------
template <int M, class PL, int NF, class S>
__global__ __launch_bounds__(256) void k_p3_fwd(EngineArgs a)
{
    constexpr int V = M / kWave;
    constexpr int HP = spec_pitch(M);
    __shared__ float2 lds[4][line_slots(M)];
    const int lane = lane_id(), w = wave_id();
    const int slot = blockIdx.y;
    const Tile t = a.tiles[slot];
    const int xi = blockIdx.x * 4 + w;
    if (xi >= t.ox1 - t.ox0) return;
    PL FP;
    FP.load(a.tw, lane);
    const int N = a.N, e = a.e, W = a.W;
    const int gx = t.ox0 + xi;
    float2* line = lds[w];
    // this column of X2 plane p
    auto col = [&](int p) {
        return plane_at<const S>(a.X2, (long long)slot * a.np2 + p, (long long)a.S2 * HP) +
               (long long)xi * HP;
    };
    // output rows [oy0, oy1) and the column lie inside the valid region, so
    // every gradient stencil is inside the image
    const Col3 F0 = col3(a.f0 + (long long)t.b * a.img_stride, N, gx);
    const long long ob = (long long)t.b * a.w_stride + (long long)(gx - e) * W - e;
    float* out = a.out0 + ob;  // indexed by the global row

    // u(g) for this column at tile rows y = lane + 64 j
    float2 u0[V];
    ifft_pair<M, PL>(FP, line, col(0), col(1), u0, lane);
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
        ifft_pair<M, PL>(FP, line, col(2), col(3), u1, lane);
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

template <int M, class PL, int NF, class SL>
__global__ __launch_bounds__(P3A_WAVES * kWave)
__attribute__((amdgpu_waves_per_eu(M <= 512 && NF == 1 && !IsMfmaPlan<PL>::value ? 4 : 1))) void k_p3_adj(EngineArgs a, int with_u, int planeL)
{
    constexpr int nf = NF;
    constexpr int V = M / kWave;
    constexpr int HP = spec_pitch(M);
    constexpr int CBK = 2 * P3A_WAVES;  // output columns per block
    __shared__ float2 lds[P3A_WAVES][line_slots(M)];
    __shared__ float wx[(CBK + 2) * M];  // w_x of block columns -1 .. CBK, rows y
    const int lane = lane_id(), w = wave_id();
    const int slot = blockIdx.y;
    const Tile t = a.tiles[slot];
    const int jb = blockIdx.x * CBK;
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
        return xi2 < ncols ? plane_at<const T>(a.X2, (long long)slot * a.np2 + plane, pstride) +
                                 (long long)xi2 * HP
                           : reinterpret_cast<const T*>(a.zcol);
    };
    // u planes: stored only for global columns [e, N - e) (u_columns)
    auto col = [&](int plane, int j) {
        const int c = t.ox0 + j;
        return c >= a.e && c < a.N - a.e ? colS(float2{}, plane, j)
                                         : static_cast<const float2*>(a.zcol);
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
                    ifft_pair<M, PL>(FP, line, col(px, j), col(py, j), u, lane);
                } else {
#pragma unroll
                    for (int q = 0; q < V; ++q) u[q] = make_float2(0.f, 0.f);
                }
            };
            cot(gx, cv);
            u_pair(j0);
#pragma unroll
            for (int j = 0; j < V; ++j) {
                wx[(2 * w + 1) * M + lane + kWave * j] = u[j].x * cv[j];
                wy0[j] = u[j].y * cv[j];
            }
            cot(gx + 1, cv);
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
                    ifft_pair<M, PL>(FP, line, col(px, jb - 1), col(px, jb + CBK), u, lane);
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
        ifft_pair<M, PL>(FP, line, colS(SL{}, planeL, j0), colS(SL{}, planeL, j0 + 1), l, lane);
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
        const dim3 g1(M / 8, c), g1a(M / 4, c);
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
            const dim3 g2(P2Items<M, RY, p2c_items<M, PL>()>::blocks(c, nf));
            run(kP2Conv, ls, [&] {
                if (x16)
                    k_p2_conv<M, RY, NW, PL, SX><<<g2, b2, lds2c, ls>>>(a, nf, c);
                else
                    k_p2_conv<M, RY, NW, PL, float2><<<g2, b2, lds2c, ls>>>(a, nf, c);
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
