// srk_api.cpp -- C ABI of the SRK shrinkage library: workspaces, tile tables,
// FP64 kernel spectra, and the drop-in entry points declared in include/srk.h.
#include "srk.h"

#include "srk_engine.hpp"
#include "srk_cd.hpp"
#include "srk_spectra.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

namespace {

thread_local int g_last_hip = 0;

int hip_status(hipError_t e)
{
    if (e != hipSuccess) {
        g_last_hip = (int)e;
        return SRK_ERR_HIP;
    }
    return SRK_OK;
}
int hip_status(int e) { return hip_status((hipError_t)e); }

bool supported_M(int M) { return M == 256 || M == 512 || M == 1024 || M == 2048; }

template <typename T>
bool aligned(const T* p, size_t a)
{
    return (reinterpret_cast<uintptr_t>(p) % a) == 0;
}

// ---------------------------------------------------------------------------
// FP64 modified Bessel K1.  x <= 2: ascending series A&S 9.6.11 (n = 1);
// x > 2: K1(x) = int_0^inf exp(-x cosh t) cosh t dt (A&S 9.6.24) by the
// trapezoid rule, which converges geometrically for this entire integrand.
// Replaces the FP32 A&S 9.8.7/9.8.8 fits of tan_tanadj/main_tan_tanadj.cu:57-89.
// ---------------------------------------------------------------------------
double k1_series(double x)
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
    return 1.0 / x + std::log(0.5 * x) * i1 - 0.25 * x * s2;
}

double k1_integral(double x)
{
    const double h = 0.0625;
    double acc = 0.5 * std::exp(-x);
    for (int k = 1; k < 20000; ++k) {
        const double c = std::cosh(k * h);
        const double v = std::exp(-x * c) * c;
        acc += v;
        if (v < 1e-20 * acc) break;
    }
    return h * acc;
}

double bessel_k1(double x) { return x <= 2.0 ? k1_series(x) : k1_integral(x); }

// SRSRK Green's tensor in FP64 (training/main_cd_training.cu:63-77,
// tan_tanadj/main_tan_tanadj.cu:93-116): K[i*kw + j], i = x (column) index.
void srsrk_kernel(int kw, double sigma, double dx, double cw, std::vector<double>& Kx,
                  std::vector<double>& Ky)
{
    const double gamma = 1.0 / sigma, Rw = cw * sigma, coeff = gamma / (2.0 * M_PI);
    const double eps = 1e-6 * dx;
    const int c = kw / 2;
    Kx.assign((size_t)kw * kw, 0.0);
    Ky.assign((size_t)kw * kw, 0.0);
    for (int i = 0; i < kw; ++i)
        for (int j = 0; j < kw; ++j) {
            const double x = (i - c) * dx, y = (j - c) * dx;
            const double r = std::sqrt(x * x + y * y);
            if (r < eps || r > Rw) continue;
            const double t = r / Rw;
            const double win = 0.42 + 0.5 * std::cos(M_PI * t) + 0.08 * std::cos(2.0 * M_PI * t);
            const double base = coeff * bessel_k1(gamma * r) * win;
            Kx[(size_t)i * kw + j] = base * (x / r);
            Ky[(size_t)i * kw + j] = base * (y / r);
        }
}

// 2-D DFT of a kw x kh col-major kernel centred at offset 0 (the reference's
// pad + fftshift, srk_forward.cuh:162-187), in double, written in the
// reference spectrum layout out[k_x*(M/2+1) + k_y].  Separable evaluation.
void kernel_dft(const std::vector<double>& G, int kw, int kh, int M, std::vector<float2>& out)
{
    const int H = M / 2 + 1, cx = kw / 2, cy = kh / 2;
    std::vector<std::complex<double>> tw(M);
    for (int n = 0; n < M; ++n) tw[n] = std::polar(1.0, -2.0 * M_PI * n / M);
    auto wmod = [M](long long v) { return (int)(((v % M) + M) % M); };
    std::vector<std::complex<double>> T((size_t)kw * H);
    for (int i = 0; i < kw; ++i)
        for (int ky = 0; ky < H; ++ky) {
            std::complex<double> acc = 0.0;
            for (int j = 0; j < kh; ++j) {
                const double g = G[(size_t)i * kh + j];
                if (g != 0.0) acc += g * tw[wmod((long long)ky * (j - cy))];
            }
            T[(size_t)i * H + ky] = acc;
        }
    out.assign((size_t)M * H, make_float2(0.f, 0.f));
    const int nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int p = 0; p < nth; ++p)
        th.emplace_back([&, p]() {
            for (int kx = p; kx < M; kx += nth)
                for (int ky = 0; ky < H; ++ky) {
                    std::complex<double> acc = 0.0;
                    for (int i = 0; i < kw; ++i)
                        acc += tw[wmod((long long)kx * (i - cx))] * T[(size_t)i * H + ky];
                    out[(size_t)kx * H + ky] = make_float2((float)acc.real(), (float)acc.imag());
                }
        });
    for (auto& t : th) t.join();
}

struct SpecKey {
    int kw;
    double sigma, dx, cw;
    int M;
    bool operator<(const SpecKey& o) const
    {
        return std::tie(kw, sigma, dx, cw, M) < std::tie(o.kw, o.sigma, o.dx, o.cw, o.M);
    }
};
struct SpecVal {
    std::vector<float2> gx, gy;
    std::vector<float> kx, ky;
};
std::mutex g_spec_mu;
std::map<SpecKey, SpecVal> g_spec_cache;

// ---------------------------------------------------------------------------
// tile tables
// ---------------------------------------------------------------------------
// Forward: outputs cover the valid region [e, N-e)^2.  N <= M: one window at
// the origin (circular wrap cannot reach the valid region because e >= KW/2).
// Otherwise windows of stride M - 2e at p0 = t*S (the reference's tiling,
// srk_forward.cuh:378-421).
bool fwd_tiles(int N, int e, int M, std::vector<srk::Tile>& t, int& S2)
{
    const int W = N - 2 * e;
    t.clear();
    if (N <= M) {
        t.push_back({0, 0, 0, e, N - e, e, N - e, 0});
        S2 = W;
        return true;
    }
    const int S = M - 2 * e;
    if (S < 1) return false;
    const int T = (W + S - 1) / S;
    for (int ty = 0; ty < T; ++ty)
        for (int tx = 0; tx < T; ++tx) {
            const int ox0 = e + tx * S, oy0 = e + ty * S;
            t.push_back({0, tx * S, ty * S, ox0, std::min(ox0 + S, N - e), oy0,
                         std::min(oy0 + S, N - e), 0});
        }
    S2 = S;
    return true;
}

// Adjoint: outputs cover all of [0, N)^2 (gather form, no overlap-add).
// N <= M: one window at the origin; the sources vanish outside [e, N-e), so
// wrapped taps read zeros.  Otherwise stride S = M - 2e - 2 with windows at
// t*S - e - 1, so u is exact one pixel beyond each output tile (for the
// stencil) and lambda_g is exact on the tile.
bool adj_tiles(int N, int e, int M, std::vector<srk::Tile>& t, int& S2)
{
    t.clear();
    if (N <= M) {
        t.push_back({0, 0, 0, 0, N, 0, N, 0});
        S2 = N + 2;
        return true;
    }
    const int S = M - 2 * e - 2;
    if (S < 1) return false;
    const int T = (N + S - 1) / S;
    for (int ty = 0; ty < T; ++ty)
        for (int tx = 0; tx < T; ++tx) {
            const int ox0 = tx * S, oy0 = ty * S;
            t.push_back({0, ox0 - e - 1, oy0 - e - 1, ox0, std::min(ox0 + S, N), oy0,
                         std::min(oy0 + S, N), 0});
        }
    S2 = S + 2;
    return true;
}

}  // namespace

static_assert(SRK_KERNEL_KINDS == srk::kKernelKinds, "kernel kinds of the C ABI and the engine");

namespace srk {

void note_hip_error(int e) { g_last_hip = e; }

hipError_t LaunchTimer::record(hipStream_t s, hipEvent_t* ev)
{
    if (used == pool.size()) {
        hipEvent_t e = nullptr;
        const hipError_t st = hipEventCreate(&e);
        if (st != hipSuccess) return st;
        pool.push_back(e);
    }
    *ev = pool[used++];
    return hipEventRecord(*ev, s);
}

hipError_t LaunchTimer::collect(double* ms, int* launches, long long* tiles)
{
    hipError_t st = hipSuccess;
    for (const Rec& r : recs) {
        float t = 0.f;
        if (st == hipSuccess) st = hipEventSynchronize(r.e1);
        if (st == hipSuccess) st = hipEventElapsedTime(&t, r.e0, r.e1);
        if (st != hipSuccess || r.kind < 0 || r.kind >= kKernelKinds) continue;
        if (ms) ms[r.kind] += t;
        if (launches) launches[r.kind] += 1;
        if (tiles) tiles[r.kind] += r.tiles;
    }
    recs.clear();
    used = 0;
    return st;
}

LaunchTimer::~LaunchTimer()
{
    for (hipEvent_t e : pool) hipEventDestroy(e);
}

}  // namespace srk

struct srk_workspace_s {
    int N = 0, W = 0, e = 0, M = 0, B_max = 0;
    int chunk_override = 0;
    int p2_rows = 0;
    int precision = SRK_PREC_FP32;
    bool timing = false;
    srk::LaunchTimer timer;
    bool concurrent = true;  // chunks alternate between two lanes (streams + buffer halves)
    srk::Fork fork{nullptr, nullptr, nullptr, nullptr, nullptr};
    srk::Tile* d_ftiles = nullptr;
    srk::Tile* d_atiles = nullptr;
    int src_b0 = 0, src_b1 = 0;  // P1A column blocks with valid-region columns (EngineArgs)
    int nft = 0, nat = 0, S2f = 0, S2a = 0;
    float2* X1 = nullptr;
    float2* X2 = nullptr;
    size_t x1_bytes = 0, x2_bytes = 0;
    float2* GT = nullptr;
    float2* tw = nullptr;
    // spectra handed over at another size (srkTrainingForward with N != M,
    // srkCheckSpectraSupport): FP64 scratch and the re-transformed spectra
    void* rs_scratch = nullptr;
    float2* rs_g = nullptr;  // [2][M][M/2+1], reference layout
    int rs_ns_max = 0;
    // fused pipeline (N <= M, M = 256 or 512): exchange rings of 16 slots and
    // the queue counters; used for FP32 forward / adjoint calls when enabled
    float2* X1f = nullptr;
    float2* X2f = nullptr;
    void* fsync = nullptr;
    bool fused = false;
    int fused_slots = 4, fused_lag = 2;  // ring slots per XCD, stage lag in rounds
    // completion of the last call (recorded on its stream): a call on another
    // stream waits for it, since calls share the exchange buffers and spectra
    hipEvent_t done = nullptr;
    hipStream_t last = nullptr;
    bool has_last = false;
    size_t total = 0;
};

namespace {

void ws_free(srk_workspace_s* w)
{
    if (!w) return;
    if (w->done) hipEventDestroy(w->done);
    if (w->fork.join) hipEventDestroy(w->fork.join);
    if (w->fork.fork) hipEventDestroy(w->fork.fork);
    if (w->fork.aux) hipStreamDestroy(w->fork.aux);
    hipFree(w->d_ftiles);
    hipFree(w->d_atiles);
    hipFree(w->X1);
    hipFree(w->X2);
    hipFree(w->GT);
    hipFree(w->tw);
    hipFree(w->rs_scratch);
    hipFree(w->rs_g);
    hipFree(w->X1f);
    hipFree(w->X2f);
    hipFree(w->fsync);
    delete w;
}

// Tiles of one lane's exchange buffers for an operator with np1 / np2 planes.
size_t lane_capacity(const srk_workspace_s* w, int np1, int np2, int S2, int lanes)
{
    const int HP = srk::spec_pitch(w->M);
    const size_t per1 = (size_t)np1 * w->M * HP * sizeof(float2);
    const size_t per2 = (size_t)np2 * S2 * HP * sizeof(float2);
    return std::min(w->x1_bytes / lanes / per1, w->x2_bytes / lanes / per2);
}

// Two lanes (each with half of the exchange buffers) when pipelining is on,
// the call's ntiles tiles take at least sixteen chunks of the whole buffers
// (or of the chunk the caller set) and half of the buffers holds at least one
// tile; otherwise one.  Split over two lanes, a call pays for twice the
// launches, smaller chunks and the fork and join: a two-image adjoint at
// N = 512 took 125 against 51 us, the N = 16384 fwd+adj step (two and three
// chunks per call) 9.50 against 9.40 ms, and the N = 512, B = 5000 step
// (forward in 3.6 chunks, adjoint in 8.8) 27.73 against 27.66 ms; at
// B = 10000 (7.1 and 17.6 chunks) the overlap wins, 55.37 against 55.42 ms.
int lane_count(const srk_workspace_s* w, int np1, int np2, int S2, int ntiles)
{
    size_t one = lane_capacity(w, np1, np2, S2, 1);
    if (w->chunk_override > 0) one = std::min<size_t>(one, (size_t)w->chunk_override);
    return w->concurrent && (size_t)ntiles > 15 * one && lane_capacity(w, np1, np2, S2, 2) >= 1
               ? 2
               : 1;
}

int chunk_for(const srk_workspace_s* w, int np1, int np2, int S2, int ntiles)
{
    const size_t cap = lane_capacity(w, np1, np2, S2, lane_count(w, np1, np2, S2, ntiles));
    size_t c = w->chunk_override > 0 ? std::min<size_t>(w->chunk_override, cap) : cap;
    c = std::max<size_t>(1, std::min<size_t>(c, (size_t)ntiles));
    return (int)c;
}

srk::EngineArgs base_args(const srk_workspace_s* w, float dx)
{
    (void)dx;
    srk::EngineArgs a{};
    a.N = w->N;
    a.e = w->e;
    a.W = w->W;
    a.img_stride = (long long)w->N * w->N;
    a.w_stride = (long long)w->W * w->W;
    a.GT = w->GT;
    a.tw = w->tw;
    a.zcol = w->tw + w->M + 512;
    a.X1 = w->X1;
    a.X2 = w->X2;
    a.p2_rows = w->p2_rows;
    return a;
}

int check_ws(const srk_workspace_s* w, int N, int W, int e, int B)
{
    if (!w) return SRK_ERR_NULL_PTR;
    if (W != N - 2 * e || W <= 0 || e < 1) return SRK_ERR_BAD_DIMS;
    if (N != w->N || W != w->W || e != w->e) return SRK_ERR_WORKSPACE;
    if (B < 1 || B > w->B_max) return SRK_ERR_BATCH;
    return SRK_OK;
}

int check_f(const float* p)
{
    if (!p) return SRK_ERR_NULL_PTR;
    if (!aligned(p, sizeof(float))) return SRK_ERR_MISALIGNED;
    return SRK_OK;
}

int check_spec(const void* gx, const void* gy)
{
    if (!gx || !gy) return SRK_ERR_NULL_PTR;
    if (!aligned((const char*)gx, 8) || !aligned((const char*)gy, 8)) return SRK_ERR_MISALIGNED;
    return SRK_OK;
}

#define SRK_TRY(expr)                 \
    do {                              \
        int _s = (expr);              \
        if (_s != SRK_OK) return _s;  \
    } while (0)

bool capturing(hipStream_t s)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
}

// Calls on one workspace share its exchange buffers, spectra and scratch.  A
// call on a stream other than the previous call's waits for that call's
// completion event first, and every call records the event at its end, so
// calls issued from one host thread on several streams do not race.  Calls
// being captured into a graph are left out (their order is the graph's).
int ws_order(srk_workspace_s* w, hipStream_t s)
{
    if (capturing(s)) return SRK_OK;
    if (w->has_last && w->last != s) SRK_TRY(hip_status(hipStreamWaitEvent(s, w->done, 0)));
    return SRK_OK;
}

int ws_record(srk_workspace_s* w, hipStream_t s)
{
    if (capturing(s)) return SRK_OK;
    SRK_TRY(hip_status(hipEventRecord(w->done, s)));
    w->last = s;
    w->has_last = true;
    return SRK_OK;
}

int import_spectra(srk_workspace_s* w, const void* gx, const void* gy, float dx, hipStream_t s)
{
    SRK_TRY(ws_order(w, s));
    const float scale = (float)((double)dx * dx / ((double)w->M * w->M));
    return hip_status(srk::engine_import_spectra(w->M, (const float2*)gx, (const float2*)gy,
                                                 scale, w->GT, s));
}

// Spectra of size Ns (reference layout) -> the kernel on the erosion box ->
// spectra at the workspace's tile size in w->rs_g (when `resample`).  Unless
// the stream is being captured, waits for the support measure and returns
// SRK_ERR_KERNEL when the Ns-periodic kernel reaches beyond the box |x|, |y|
// <= erosion (relative RMS outside above 1e-4): the valid region would then
// see taps the erosion does not cover.  *ratio_out (optional) receives that
// measure (-1 while capturing).
constexpr double kSupportTol = 1e-4;

int convert_spectra(srk_workspace_s* w, const void* gx, const void* gy, int Ns, bool resample,
                    hipStream_t s, double* ratio_out)
{
    if (Ns < 2 * w->e + 1 || Ns > w->rs_ns_max) return SRK_ERR_UNSUPPORTED_SIZE;
    SRK_TRY(ws_order(w, s));
    double* d_ratio = nullptr;
    float2* dst = resample ? w->rs_g : nullptr;
    const size_t plane = (size_t)w->M * (w->M / 2 + 1);
    SRK_TRY(hip_status(srk::resample_spectra_device(
        Ns, w->e, w->M, (const float2*)gx, (const float2*)gy, w->rs_scratch, dst,
        dst ? dst + plane : nullptr, &d_ratio, w->rs_ns_max, s)));
    if (ratio_out) *ratio_out = -1.0;
    if (capturing(s)) return SRK_OK;
    SRK_TRY(ws_record(w, s));
    double r = 0.0;
    SRK_TRY(hip_status(hipMemcpyAsync(&r, d_ratio, sizeof(double), hipMemcpyDeviceToHost, s)));
    SRK_TRY(hip_status(hipStreamSynchronize(s)));
    if (ratio_out) *ratio_out = r;
    return r <= kSupportTol ? SRK_OK : SRK_ERR_KERNEL;
}

// The launch timer to use on stream s: none unless timing is on and s is not
// being captured (event records would become graph nodes).
srk::LaunchTimer* timer_for(srk_workspace_s* w, hipStream_t s)
{
    if (!w->timing) return nullptr;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
        return nullptr;
    return &w->timer;
}

// The second lane (stream, event pair and the upper halves of the exchange
// buffers), or none.
const srk::Fork* lanes_for(srk_workspace_s* w, const srk::EngineArgs& a, int ntiles)
{
    if (lane_count(w, a.np1, a.np2, a.S2, ntiles) < 2) return nullptr;
    w->fork.X1b = w->X1 + w->x1_bytes / 2 / sizeof(float2);
    w->fork.X2b = w->X2 + w->x2_bytes / 2 / sizeof(float2);
    return &w->fork;
}

// Exchange rings of the fused pipeline, allocated when it is first enabled:
// kFusedSlotsMax slots per XCD of up to 3 X1 planes (adjoint: g, a_x, a_y)
// and 3 X2 planes of max(W, N + 2) columns, and the queue counters.
int fused_buffers(srk_workspace_s* w)
{
    if (w->X1f || !srk::fused_supported(srk::Op::Forward, w->M, w->N, 0)) return SRK_OK;
    const int M = w->M, HP = srk::spec_pitch(M);
    const size_t slots = (size_t)8 * srk::kFusedSlotsMax;
    const size_t b1 = slots * 3 * M * HP * sizeof(float2);
    const size_t b2 = slots * 3 * std::max(w->S2f, w->S2a) * HP * sizeof(float2);
    const size_t bs = srk::fused_sync_bytes();
    float2 *x1 = nullptr, *x2 = nullptr;
    void* sy = nullptr;
    hipError_t e = hipMalloc((void**)&x1, b1);
    if (e == hipSuccess) e = hipMalloc((void**)&x2, b2);
    if (e == hipSuccess) e = hipMalloc(&sy, bs);
    if (e == hipSuccess) e = hipMemset(sy, 0, bs);
    if (e != hipSuccess) {
        hipFree(x1);
        hipFree(x2);
        hipFree(sy);
        return hip_status(e);
    }
    w->X1f = x1;
    w->X2f = x2;
    w->fsync = sy;
    w->total += b1 + b2 + bs;
    return SRK_OK;
}

// the fused pipeline serves this call
bool use_fused(const srk_workspace_s* w, srk::Op op)
{
    return w->fused && w->X1f && srk::fused_supported(op, w->M, w->N, w->precision);
}

int run_fwd(srk_workspace_s* w, srk::Op op, srk::EngineArgs a, int B, hipStream_t s)
{
    const int nf = op == srk::Op::Tangent ? 2 : 1;
    a.tiles = w->d_ftiles;
    a.np1 = nf;
    a.np2 = 2 * nf;
    a.S2 = w->S2f;
    a.halo2 = 0;
    if (use_fused(w, op)) {
        a.X1 = w->X1f;
        a.X2 = w->X2f;
        SRK_TRY(hip_status(srk::engine_run_fused(op, w->M, a, B, w->fsync, w->fused_slots,
                                                 w->fused_lag, s, timer_for(w, s))));
        return ws_record(w, s);
    }
    const int ntiles = B * w->nft;
    const int chunk = chunk_for(w, a.np1, a.np2, a.S2, ntiles);
    SRK_TRY(hip_status(srk::engine_run(op, w->M, w->precision, a, ntiles, chunk, s,
                                       timer_for(w, s), lanes_for(w, a, ntiles))));
    return ws_record(w, s);
}

int run_adj(srk_workspace_s* w, srk::Op op, srk::EngineArgs a, int B, hipStream_t s)
{
    switch (op) {
    case srk::Op::Adjoint: a.np1 = 3; a.np2 = 3; break;
    case srk::Op::AdjointG: a.np1 = 2; a.np2 = 1; break;
    case srk::Op::TangentAdjoint: a.np1 = 4; a.np2 = 5; break;
    default: return SRK_ERR_INVALID_ARG;
    }
    a.tiles = w->d_atiles;
    a.S2 = w->S2a;
    a.halo2 = 1;
    a.src_b0 = w->src_b0;
    a.src_b1 = w->src_b1;
    if (use_fused(w, op)) {
        a.X1 = w->X1f;
        a.X2 = w->X2f;
        SRK_TRY(hip_status(srk::engine_run_fused(op, w->M, a, B, w->fsync, w->fused_slots,
                                                 w->fused_lag, s, timer_for(w, s))));
        return ws_record(w, s);
    }
    const int ntiles = B * w->nat;
    const int chunk = chunk_for(w, a.np1, a.np2, a.S2, ntiles);
    SRK_TRY(hip_status(srk::engine_run(op, w->M, w->precision, a, ntiles, chunk, s,
                                       timer_for(w, s), lanes_for(w, a, ntiles))));
    return ws_record(w, s);
}

}  // namespace

extern "C" {

const char* srkGetErrorString(int status)
{
    switch (status) {
    case SRK_OK: return "success";
    case SRK_ERR_INVALID_ARG: return "invalid argument";
    case SRK_ERR_BAD_DIMS: return "inconsistent dimensions (need W = N - 2*erosion > 0, erosion >= 1)";
    case SRK_ERR_NULL_PTR: return "null pointer";
    case SRK_ERR_MISALIGNED: return "misaligned pointer";
    case SRK_ERR_UNSUPPORTED_SIZE: return "unsupported tile size M (256, 512, 1024, 2048; M > 2*erosion + 2)";
    case SRK_ERR_WORKSPACE: return "workspace was created for different N / W / erosion / M";
    case SRK_ERR_BATCH: return "batch size outside [1, B_max]";
    case SRK_ERR_HIP: return "HIP runtime error";
    case SRK_ERR_KERNEL: return "invalid kernel parameters";
    default: return "unknown status";
    }
}

int srkLastHipError(void) { return g_last_hip; }
int srkGetVersion(void) { return 100; }

int srkSynchronize(hipStream_t stream) { return hip_status(hipStreamSynchronize(stream)); }

int srkPrecomputeKernelFreq(const float* h_G, int kw, int kh, int M, void* d_G_freq,
                            hipStream_t stream)
{
    if (!h_G || !d_G_freq) return SRK_ERR_NULL_PTR;
    if (kw < 1 || kh < 1 || kw % 2 == 0 || kh % 2 == 0 || kw > M || kh > M) return SRK_ERR_KERNEL;
    if (M < 2) return SRK_ERR_UNSUPPORTED_SIZE;  // any size, as cuFFT plans any N
    std::vector<double> G((size_t)kw * kh);
    for (size_t i = 0; i < G.size(); ++i) G[i] = h_G[i];
    std::vector<float2> out;
    kernel_dft(G, kw, kh, M, out);
    SRK_TRY(hip_status(hipMemcpyAsync(d_G_freq, out.data(), out.size() * sizeof(float2),
                                      hipMemcpyHostToDevice, stream)));
    return hip_status(hipStreamSynchronize(stream));
}

int srkKernelSpectraSRSRK(int kw, double sigma, double dx, double cw, int M, void* d_Gx_freq,
                          void* d_Gy_freq, float* h_Gx, float* h_Gy, hipStream_t stream)
{
    if (kw < 3 || kw % 2 == 0 || !(sigma > 0) || !(dx > 0) || !(cw > 0)) return SRK_ERR_KERNEL;
    if (M < 2 || kw > M) return SRK_ERR_UNSUPPORTED_SIZE;
    const SpecKey key{kw, sigma, dx, cw, M};
    SpecVal* val = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_spec_mu);
        auto it = g_spec_cache.find(key);
        if (it == g_spec_cache.end()) {
            std::vector<double> Kx, Ky;
            srsrk_kernel(kw, sigma, dx, cw, Kx, Ky);
            SpecVal v;
            kernel_dft(Kx, kw, kw, M, v.gx);
            kernel_dft(Ky, kw, kw, M, v.gy);
            v.kx.assign(Kx.begin(), Kx.end());
            v.ky.assign(Ky.begin(), Ky.end());
            it = g_spec_cache.emplace(key, std::move(v)).first;
        }
        val = &it->second;
    }
    if (h_Gx) std::copy(val->kx.begin(), val->kx.end(), h_Gx);
    if (h_Gy) std::copy(val->ky.begin(), val->ky.end(), h_Gy);
    if (d_Gx_freq)
        SRK_TRY(hip_status(hipMemcpyAsync(d_Gx_freq, val->gx.data(), val->gx.size() * sizeof(float2),
                                          hipMemcpyHostToDevice, stream)));
    if (d_Gy_freq)
        SRK_TRY(hip_status(hipMemcpyAsync(d_Gy_freq, val->gy.data(), val->gy.size() * sizeof(float2),
                                          hipMemcpyHostToDevice, stream)));
    if (!d_Gx_freq && !d_Gy_freq) return SRK_OK;  // host kernels only: no device work
    return hip_status(hipStreamSynchronize(stream));
}

int srkKernelSpectraSRSRKDevice(int kw, double sigma, double dx, double cw, int M,
                                void* d_Gx_freq, void* d_Gy_freq, float* d_Kx, float* d_Ky,
                                hipStream_t stream)
{
    if (kw < 3 || kw % 2 == 0 || !(sigma > 0) || !(dx > 0) || !(cw > 0)) return SRK_ERR_KERNEL;
    if (M < 2 || kw > M) return SRK_ERR_UNSUPPORTED_SIZE;
    if (!d_Gx_freq || !d_Gy_freq) return SRK_ERR_NULL_PTR;
    if (!aligned((const char*)d_Gx_freq, 8) || !aligned((const char*)d_Gy_freq, 8))
        return SRK_ERR_MISALIGNED;
    void* scratch = nullptr;
    SRK_TRY(hip_status(hipMallocAsync(&scratch, srk::spectra_scratch_bytes(kw, M), stream)));
    const int st = hip_status(srk::spectra_srsrk_device(kw, sigma, dx, cw, M, scratch,
                                                        (float2*)d_Gx_freq, (float2*)d_Gy_freq,
                                                        d_Kx, d_Ky, stream));
    const int st2 = hip_status(hipFreeAsync(scratch, stream));
    return st != SRK_OK ? st : st2;
}

int srkWorkspaceCreate(int N, int W, int erosion, int M, int B_max, size_t scratch_mb,
                       srk_workspace_t* out)
{
    if (!out) return SRK_ERR_NULL_PTR;
    *out = nullptr;
    if (W != N - 2 * erosion || W <= 0 || erosion < 1) return SRK_ERR_BAD_DIMS;
    if (B_max < 1) return SRK_ERR_BATCH;
    if (!supported_M(M) || M <= 2 * erosion + 2) return SRK_ERR_UNSUPPORTED_SIZE;
    std::vector<srk::Tile> ft, at;
    int S2f = 0, S2a = 0;
    if (!fwd_tiles(N, erosion, M, ft, S2f) || !adj_tiles(N, erosion, M, at, S2a))
        return SRK_ERR_UNSUPPORTED_SIZE;
    auto* w = new srk_workspace_s();
    w->N = N; w->W = W; w->e = erosion; w->M = M; w->B_max = B_max;
    w->nft = (int)ft.size(); w->nat = (int)at.size(); w->S2f = S2f; w->S2a = S2a;
    // P1A column blocks of 4 tile columns that meet the valid region [e, N - e)
    // in some adjoint tile
    w->src_b0 = M / 4;
    w->src_b1 = 0;
    for (const srk::Tile& t : at) {
        const int lo = std::max(erosion - t.p0x, 0), hi = std::min(N - erosion - t.p0x, M);
        if (hi <= lo) continue;
        w->src_b0 = std::min(w->src_b0, lo / 4);
        w->src_b1 = std::max(w->src_b1, (hi + 3) / 4);
    }
    // replicate the per-image tables over the batch
    std::vector<srk::Tile> fb((size_t)B_max * ft.size()), ab((size_t)B_max * at.size());
    for (int b = 0; b < B_max; ++b) {
        for (size_t i = 0; i < ft.size(); ++i) { fb[b * ft.size() + i] = ft[i]; fb[b * ft.size() + i].b = b; }
        for (size_t i = 0; i < at.size(); ++i) { ab[b * at.size() + i] = at[i]; ab[b * at.size() + i].b = b; }
    }
    const int HP = srk::spec_pitch(M);
    const size_t x1_per = (size_t)4 * M * HP * sizeof(float2);
    const size_t x2_per = (size_t)std::max(4 * S2f, 5 * S2a) * HP * sizeof(float2);
    // default exchange budget: 4 GiB, raised up to 16 GiB to hold every tile
    // of one image when a single image needs more (at N = 16384, M = 1024 a
    // call then runs in one chunk instead of two or three: fwd+adj 9.07
    // against 9.19 ms)
    const size_t tiles1 = (size_t)std::max(w->nft, w->nat);
    const size_t budget =
        scratch_mb ? scratch_mb << 20
                   : std::clamp<size_t>(tiles1 * (x1_per + x2_per), size_t(4096) << 20,
                                        size_t(16384) << 20);
    size_t cmax = std::max<size_t>(1, budget / (x1_per + x2_per));
    cmax = std::min<size_t>(cmax, (size_t)B_max * std::max(w->nft, w->nat));
    w->x1_bytes = cmax * x1_per;
    w->x2_bytes = cmax * x2_per;
    // twiddles exp(-2 pi i k / M), followed by the per-lane FP16 hi/lo tiles of
    // the 16-point DFT used by the matrix-core transforms (256 x 8 halves), and
    // one all-zero column spectrum (the stand-in for columns past a tile's end).
    std::vector<float2> tw(M + 512 + srk::spec_pitch(M), make_float2(0.f, 0.f));
    for (int k = 0; k < M; ++k) {
        const double a = -2.0 * M_PI * k / M;
        tw[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    {
        _Float16* tiles = reinterpret_cast<_Float16*>(tw.data() + M);
        for (int lane = 0; lane < 64; ++lane)
            for (int p = 0; p < 2; ++p)
                for (int nb = 0; nb < 2; ++nb)
                    for (int e = 0; e < 8; ++e) {
                        const int n = lane & 15, q = lane >> 4;
                        const int r = 8 * (q & 1) + e;
                        const bool in_im = q >= 2, out_im = nb == 1;
                        const double th = 2.0 * M_PI * ((r * n) & 15) / 16.0;
                        // Re X = sum a cos + b sin ; Im X = sum b cos - a sin
                        const double v = in_im ? (out_im ? std::cos(th) : std::sin(th))
                                               : (out_im ? -std::sin(th) : std::cos(th));
                        const _Float16 hi = (_Float16)v;
                        const _Float16 lo = (_Float16)(v - (double)hi);
                        tiles[(((lane * 2 + p) * 2 + nb) * 8) + e] = p == 0 ? hi : lo;
                    }
    }
    int st = SRK_OK;
    auto alloc = [&](void** p, size_t bytes) {
        if (st != SRK_OK) return;
        st = hip_status(hipMalloc(p, bytes));
        if (st == SRK_OK) w->total += bytes;
    };
    alloc((void**)&w->d_ftiles, fb.size() * sizeof(srk::Tile));
    alloc((void**)&w->d_atiles, ab.size() * sizeof(srk::Tile));
    alloc((void**)&w->X1, w->x1_bytes);
    alloc((void**)&w->X2, w->x2_bytes);
    alloc((void**)&w->GT, (size_t)2 * (M / 2 + 1) * M * sizeof(float2));
    alloc((void**)&w->tw, tw.size() * sizeof(float2));
    w->rs_ns_max = std::max(N, M);
    alloc(&w->rs_scratch, srk::resample_scratch_bytes(w->rs_ns_max, erosion, M));
    alloc((void**)&w->rs_g, (size_t)2 * M * (M / 2 + 1) * sizeof(float2));
    if (st == SRK_OK) st = hip_status(hipStreamCreateWithFlags(&w->fork.aux, hipStreamNonBlocking));
    if (st == SRK_OK)
        st = hip_status(hipEventCreateWithFlags(&w->fork.fork, hipEventDisableTiming));
    if (st == SRK_OK)
        st = hip_status(hipEventCreateWithFlags(&w->fork.join, hipEventDisableTiming));
    if (st == SRK_OK) st = hip_status(hipEventCreateWithFlags(&w->done, hipEventDisableTiming));
    if (st == SRK_OK)
        st = hip_status(hipMemcpy(w->d_ftiles, fb.data(), fb.size() * sizeof(srk::Tile), hipMemcpyHostToDevice));
    if (st == SRK_OK)
        st = hip_status(hipMemcpy(w->d_atiles, ab.data(), ab.size() * sizeof(srk::Tile), hipMemcpyHostToDevice));
    if (st == SRK_OK)
        st = hip_status(hipMemcpy(w->tw, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
    if (st != SRK_OK) {
        ws_free(w);
        return st;
    }
    if (const char* env = std::getenv("SRK_CHUNK_TILES")) w->chunk_override = std::atoi(env);
    if (const char* env = std::getenv("SRK_P2_ROWS")) w->p2_rows = std::atoi(env);
    if (const char* env = std::getenv("SRK_FUSED")) w->fused = std::atoi(env) != 0;
    if (w->fused && (st = fused_buffers(w)) != SRK_OK) {
        ws_free(w);
        return st;
    }
    if (const char* env = std::getenv("SRK_FUSED_SLOTS")) w->fused_slots = std::atoi(env);
    if (const char* env = std::getenv("SRK_FUSED_LAG")) w->fused_lag = std::atoi(env);
    *out = w;
    return SRK_OK;
}

int srkWorkspaceDestroy(srk_workspace_t w)
{
    if (!w) return SRK_ERR_NULL_PTR;
    ws_free(w);
    return SRK_OK;
}

size_t srkWorkspaceBytes(srk_workspace_t w) { return w ? w->total : 0; }

int srkWorkspaceSetChunk(srk_workspace_t w, int tiles)
{
    if (!w) return SRK_ERR_NULL_PTR;
    if (tiles < 0) return SRK_ERR_INVALID_ARG;
    w->chunk_override = tiles;
    return SRK_OK;
}

int srkWorkspaceSetTuning(srk_workspace_t w, int key, int value)
{
    if (!w) return SRK_ERR_NULL_PTR;
    switch (key) {
    case SRK_TUNE_CHUNK_TILES:
        if (value < 0) return SRK_ERR_INVALID_ARG;
        w->chunk_override = value;
        return SRK_OK;
    case SRK_TUNE_P2_ROWS:
        if (value != 0 && value != 4 && value != 8 && value != 16) return SRK_ERR_INVALID_ARG;
        w->p2_rows = value;
        return SRK_OK;
    case SRK_TUNE_CONCURRENCY:
        if (value != 0 && value != 1) return SRK_ERR_INVALID_ARG;
        w->concurrent = value != 0;
        return SRK_OK;
    case SRK_TUNE_FUSED:
        if (value != 0 && value != 1) return SRK_ERR_INVALID_ARG;
        if (value) SRK_TRY(fused_buffers(w));
        w->fused = value != 0;
        return SRK_OK;
    case SRK_TUNE_FUSED_SLOTS:
        if (value <= w->fused_lag || value > srk::kFusedSlotsMax) return SRK_ERR_INVALID_ARG;
        w->fused_slots = value;
        return SRK_OK;
    case SRK_TUNE_FUSED_LAG:
        if (value < 1 || value >= w->fused_slots) return SRK_ERR_INVALID_ARG;
        w->fused_lag = value;
        return SRK_OK;
    default:
        return SRK_ERR_INVALID_ARG;
    }
}

int srkWorkspaceFusedStatus(srk_workspace_t w, int* available, int* timeouts)
{
    if (!w) return SRK_ERR_NULL_PTR;
    if (available) *available = srk::fused_supported(srk::Op::Forward, w->M, w->N, 0) ? 1 : 0;
    if (timeouts) *timeouts = 0;
    if (!w->fsync) return SRK_OK;
    int t = 0;
    SRK_TRY(hip_status(hipDeviceSynchronize()));
    SRK_TRY(hip_status(srk::fused_sync_status(w->fsync, &t)));
    if (timeouts) *timeouts = t;
    return t == 0 ? SRK_OK : SRK_ERR_HIP;
}

int srkWorkspaceSetTiming(srk_workspace_t w, int enable)
{
    if (!w) return SRK_ERR_NULL_PTR;
    w->timing = enable != 0;
    return SRK_OK;
}

int srkWorkspaceKernelTimes(srk_workspace_t w, double* total_ms, int* launches, long long* tiles)
{
    if (!w) return SRK_ERR_NULL_PTR;
    for (int k = 0; k < SRK_KERNEL_KINDS; ++k) {
        if (total_ms) total_ms[k] = 0.0;
        if (launches) launches[k] = 0;
        if (tiles) tiles[k] = 0;
    }
    return hip_status(w->timer.collect(total_ms, launches, tiles));
}

const char* srkKernelKindName(int kind)
{
    static const char* names[SRK_KERNEL_KINDS] = {"k_p1_cols", "k_p1_adj_src", "k_p2_conv",
                                                  "k_p2_corr", "k_p3_fwd",    "k_p3_adj",
                                                  "k_fused"};
    return kind >= 0 && kind < SRK_KERNEL_KINDS ? names[kind] : "";
}

int srkWorkspaceSetPrecision(srk_workspace_t w, int precision)
{
    if (!w) return SRK_ERR_NULL_PTR;
    if (precision == SRK_PREC_FP32 || precision == SRK_PREC_BF16_EXCHANGE) {
        w->precision = precision;
        return SRK_OK;
    }
    if (precision == SRK_PREC_MIXED_FP16) {
        if (w->M > 1024) return SRK_ERR_UNSUPPORTED_SIZE;
        w->precision = precision;
        return SRK_OK;
    }
    return SRK_ERR_INVALID_ARG;
}

int srkWorkspaceInfo(srk_workspace_t w, int* M, int* fwd_tiles, int* adj_tiles, int* chunk_fwd,
                     int* chunk_adj)
{
    if (!w) return SRK_ERR_NULL_PTR;
    if (M) *M = w->M;
    if (fwd_tiles) *fwd_tiles = w->nft;
    if (adj_tiles) *adj_tiles = w->nat;
    if (chunk_fwd) *chunk_fwd = chunk_for(w, 1, 2, w->S2f, w->B_max * w->nft);
    if (chunk_adj) *chunk_adj = chunk_for(w, 3, 3, w->S2a, w->B_max * w->nat);
    return SRK_OK;
}

int srkForwardBatched(const float* E, const float* I, const void* Gx, const void* Gy, int N, int W,
                      int erosion, float dx, int B, hipStream_t stream, float* R,
                      srk_workspace_t w)
{
    SRK_TRY(check_ws(w, N, W, erosion, B));
    SRK_TRY(check_f(E)); SRK_TRY(check_f(I)); SRK_TRY(check_f(R)); SRK_TRY(check_spec(Gx, Gy));
    SRK_TRY(import_spectra(w, Gx, Gy, dx, stream));
    srk::EngineArgs a = base_args(w, dx);
    a.g0 = E; a.f0 = I; a.out0 = R;
    return run_fwd(w, srk::Op::Forward, a, B, stream);
}

int srkCheckSpectraSupport(const void* Gx, const void* Gy, int Ns, int erosion,
                           hipStream_t stream, srk_workspace_t w, double* ratio)
{
    if (!w) return SRK_ERR_NULL_PTR;
    SRK_TRY(check_spec(Gx, Gy));
    if (erosion != w->e) return SRK_ERR_WORKSPACE;
    return convert_spectra(w, Gx, Gy, Ns, false, stream, ratio);
}

int srkForwardBatchedU(const float* E, const float* I, const void* Gx, const void* Gy, int N,
                       int W, int erosion, float dx, int B, hipStream_t stream, float* R,
                       float* ux, float* uy, srk_workspace_t w)
{
    SRK_TRY(check_ws(w, N, W, erosion, B));
    SRK_TRY(check_f(E)); SRK_TRY(check_f(I)); SRK_TRY(check_f(R)); SRK_TRY(check_spec(Gx, Gy));
    SRK_TRY(check_f(ux)); SRK_TRY(check_f(uy));
    SRK_TRY(import_spectra(w, Gx, Gy, dx, stream));
    srk::EngineArgs a = base_args(w, dx);
    a.g0 = E; a.f0 = I; a.out0 = R;
    a.ux_out = ux; a.uy_out = uy;
    return run_fwd(w, srk::Op::Forward, a, B, stream);
}

int srkAdjointBatchedU(const float* v, const float* ux, const float* uy, const float* I,
                       const void* Gx, const void* Gy, int N, int W, int erosion, float dx, int B,
                       hipStream_t stream, float* dLdI, float* dLdE, srk_workspace_t w)
{
    SRK_TRY(check_ws(w, N, W, erosion, B));
    SRK_TRY(check_f(v)); SRK_TRY(check_f(ux)); SRK_TRY(check_f(uy)); SRK_TRY(check_f(I));
    SRK_TRY(check_f(dLdI)); SRK_TRY(check_f(dLdE)); SRK_TRY(check_spec(Gx, Gy));
    SRK_TRY(hip_status(srk::engine_adjoint_f_from_u(v, ux, uy, dLdI, N, W, erosion, B,
                                                    (long long)N * N, (long long)W * W, stream)));
    SRK_TRY(import_spectra(w, Gx, Gy, dx, stream));
    srk::EngineArgs a = base_args(w, dx);
    a.f0 = I; a.v0 = v; a.out1 = dLdE;
    return run_adj(w, srk::Op::AdjointG, a, B, stream);
}

int srkTrainingForward(const float* E, const float* I, const void* Gx, const void* Gy, int N,
                       int W, int erosion, float dx, int B, hipStream_t stream, float* R,
                       srk_workspace_t w)
{
    if (w && w->M == N) return srkForwardBatched(E, I, Gx, Gy, N, W, erosion, dx, B, stream, R, w);
    // spectra at size N, tiles of size M != N: the spectra fix the spatial
    // kernel on the erosion box exactly (its support must lie inside it, as
    // for the reference's circular transform at size N to be exact on the
    // valid region); re-transform it at size M
    SRK_TRY(check_ws(w, N, W, erosion, B));
    SRK_TRY(check_f(E)); SRK_TRY(check_f(I)); SRK_TRY(check_f(R)); SRK_TRY(check_spec(Gx, Gy));
    SRK_TRY(convert_spectra(w, Gx, Gy, N, true, stream, nullptr));
    const size_t plane = (size_t)w->M * (w->M / 2 + 1);
    SRK_TRY(import_spectra(w, w->rs_g, w->rs_g + plane, dx, stream));
    srk::EngineArgs a = base_args(w, dx);
    a.g0 = E; a.f0 = I; a.out0 = R;
    return run_fwd(w, srk::Op::Forward, a, B, stream);
}

int srkForwardOaS(const float* E, const float* I, const void* Gx, const void* Gy, int N, int W,
                  int erosion, float dx, hipStream_t stream, float* R, srk_workspace_t w,
                  float* ux, float* uy)
{
    SRK_TRY(check_ws(w, N, W, erosion, 1));
    SRK_TRY(check_f(E)); SRK_TRY(check_f(I)); SRK_TRY(check_f(R)); SRK_TRY(check_spec(Gx, Gy));
    if ((ux == nullptr) != (uy == nullptr)) return SRK_ERR_NULL_PTR;
    if (ux) { SRK_TRY(check_f(ux)); SRK_TRY(check_f(uy)); }
    SRK_TRY(import_spectra(w, Gx, Gy, dx, stream));
    srk::EngineArgs a = base_args(w, dx);
    a.g0 = E; a.f0 = I; a.out0 = R;
    a.ux_out = ux; a.uy_out = uy;
    return run_fwd(w, srk::Op::Forward, a, 1, stream);
}

int adjointSRK(const float* v, const float* ux, const float* uy, const float* I, const void* Gx,
               const void* Gy, int N, int W, int erosion, float dx, hipStream_t stream,
               float* dLdI, float* dLdE, srk_workspace_t w)
{
    SRK_TRY(check_ws(w, N, W, erosion, 1));
    SRK_TRY(check_f(v)); SRK_TRY(check_f(ux)); SRK_TRY(check_f(uy)); SRK_TRY(check_f(I));
    SRK_TRY(check_f(dLdI)); SRK_TRY(check_f(dLdE)); SRK_TRY(check_spec(Gx, Gy));
    SRK_TRY(hip_status(srk::engine_adjoint_f_from_u(v, ux, uy, dLdI, N, W, erosion, 1,
                                                    (long long)N * N, (long long)W * W, stream)));
    SRK_TRY(import_spectra(w, Gx, Gy, dx, stream));
    srk::EngineArgs a = base_args(w, dx);
    a.f0 = I; a.v0 = v; a.out1 = dLdE;
    return run_adj(w, srk::Op::AdjointG, a, 1, stream);
}

int srkAdjointBatched(const float* v, const float* E, const float* I, const void* Gx,
                      const void* Gy, int N, int W, int erosion, float dx, int B,
                      hipStream_t stream, float* dLdI, float* dLdE, srk_workspace_t w)
{
    SRK_TRY(check_ws(w, N, W, erosion, B));
    SRK_TRY(check_f(v)); SRK_TRY(check_f(E)); SRK_TRY(check_f(I));
    SRK_TRY(check_f(dLdI)); SRK_TRY(check_f(dLdE)); SRK_TRY(check_spec(Gx, Gy));
    SRK_TRY(import_spectra(w, Gx, Gy, dx, stream));
    srk::EngineArgs a = base_args(w, dx);
    a.g0 = E; a.f0 = I; a.v0 = v; a.out0 = dLdI; a.out1 = dLdE;
    return run_adj(w, srk::Op::Adjoint, a, B, stream);
}

int srkTangentBatched(const float* f, const float* g, const float* df, const float* dg,
                      const void* Gx, const void* Gy, int N, int W, int erosion, float dx, int B,
                      hipStream_t stream, float* dh, srk_workspace_t w)
{
    SRK_TRY(check_ws(w, N, W, erosion, B));
    SRK_TRY(check_f(f)); SRK_TRY(check_f(g)); SRK_TRY(check_f(df)); SRK_TRY(check_f(dg));
    SRK_TRY(check_f(dh)); SRK_TRY(check_spec(Gx, Gy));
    SRK_TRY(import_spectra(w, Gx, Gy, dx, stream));
    srk::EngineArgs a = base_args(w, dx);
    a.g0 = g; a.g1 = dg; a.f0 = f; a.f1 = df; a.out0 = dh;
    return run_fwd(w, srk::Op::Tangent, a, B, stream);
}

int shrinkageTangentOaS(const float* f, const float* g, const float* df, const float* dg,
                        const void* Gx, const void* Gy, int N, int W, int erosion, float dx,
                        hipStream_t stream, float* dh, srk_workspace_t w)
{
    return srkTangentBatched(f, g, df, dg, Gx, Gy, N, W, erosion, dx, 1, stream, dh, w);
}

int srkTangentAdjointBatched(const float* f, const float* g, const float* df, const float* dg,
                             const float* lam_h, const float* dlam_h, const void* Gx,
                             const void* Gy, int N, int W, int erosion, float dx, int B,
                             hipStream_t stream, float* dLam_f, float* dLam_g,
                             srk_workspace_t w)
{
    SRK_TRY(check_ws(w, N, W, erosion, B));
    SRK_TRY(check_f(f)); SRK_TRY(check_f(g)); SRK_TRY(check_f(df)); SRK_TRY(check_f(dg));
    SRK_TRY(check_f(lam_h)); SRK_TRY(check_f(dlam_h));
    SRK_TRY(check_f(dLam_f)); SRK_TRY(check_f(dLam_g)); SRK_TRY(check_spec(Gx, Gy));
    SRK_TRY(import_spectra(w, Gx, Gy, dx, stream));
    srk::EngineArgs a = base_args(w, dx);
    // w_d = u_d(g) dlam_h + u_d(dg) lam_h ;  a_d = dlam_h d_d f + lam_h d_d df
    a.g0 = g; a.g1 = dg; a.f0 = f; a.f1 = df; a.v0 = dlam_h; a.v1 = lam_h;
    a.out0 = dLam_f; a.out1 = dLam_g;
    return run_adj(w, srk::Op::TangentAdjoint, a, B, stream);
}

int shrinkageTangentAdjointOaS(const float* f, const float* g, const float* df,
                               const float* dg, const float* lam_h, const float* dlam_h,
                               const void* Gx, const void* Gy, int N, int W, int erosion,
                               float dx, hipStream_t stream, float* dLam_f, float* dLam_g,
                               srk_workspace_t w)
{
    return srkTangentAdjointBatched(f, g, df, dg, lam_h, dlam_h, Gx, Gy, N, W, erosion, dx, 1,
                                    stream, dLam_f, dLam_g, w);
}

int srkMeasureCDBatched(const float* d_in, int B, int N, int L, int A, float threshold,
                        float* d_left_px, float* d_right_px, float* d_cd_px, hipStream_t stream)
{
    if (!d_in || !d_left_px || !d_right_px || !d_cd_px) return SRK_ERR_NULL_PTR;
    if (B < 1) return SRK_ERR_BATCH;
    if (N < 2 || L < 1 || A < 1) return SRK_ERR_BAD_DIMS;
    if (!srk::cd_shape_ok(N, L, A)) return SRK_ERR_UNSUPPORTED_SIZE;
    return hip_status(srk::cd_measure_rows(d_in, B, N, L, A, threshold, d_left_px, d_right_px,
                                           d_cd_px, stream));
}

int srkBuildPolyphaseTable(int L, int A, float* d_coeff, hipStream_t stream)
{
    if (!d_coeff) return SRK_ERR_NULL_PTR;
    if (!aligned(d_coeff, sizeof(float))) return SRK_ERR_MISALIGNED;
    if (L < 1 || A < 1 || L > 4096 || A > 256) return SRK_ERR_BAD_DIMS;
    return hip_status(srk::cd_build_table(L, A, d_coeff, stream));
}

int srkUpsampleBatched(const float* d_in, const float* d_coeff, int B, int N, int L, int A,
                       float* d_out, hipStream_t stream)
{
    if (!d_in || !d_coeff || !d_out) return SRK_ERR_NULL_PTR;
    if (!aligned(d_in, sizeof(float)) || !aligned(d_coeff, sizeof(float)) ||
        !aligned(d_out, sizeof(float)))
        return SRK_ERR_MISALIGNED;
    if (B < 1) return SRK_ERR_BATCH;
    if (N < 1 || L < 1 || A < 1 || L > 4096 || A > 256) return SRK_ERR_BAD_DIMS;
    if ((long long)N * L > (1LL << 31) - 1) return SRK_ERR_UNSUPPORTED_SIZE;
    return hip_status(srk::cd_upsample(d_in, d_coeff, B, N, L, A, d_out, stream));
}

int srkMeasureCDBatchedEx(const float* d_in, const float* d_coeff, const float* d_up, int B, int N,
                          int L, int A, float* d_left_px, float* d_right_px, float* d_cd_px,
                          hipStream_t stream)
{
    if ((!d_in && !d_up) || !d_left_px || !d_right_px || !d_cd_px) return SRK_ERR_NULL_PTR;
    if (B < 1) return SRK_ERR_BATCH;
    if (N < 2 || L < 1 || A < 1) return SRK_ERR_BAD_DIMS;
    if (d_up) {
        if ((long long)N * L > (1LL << 24)) return SRK_ERR_UNSUPPORTED_SIZE;
    } else if (!srk::cd_shape_ok(N, L, A)) {
        return SRK_ERR_UNSUPPORTED_SIZE;
    }
    return hip_status(srk::cd_measure_ex(d_in, d_coeff, d_up, B, N, L, A, d_left_px, d_right_px,
                                         d_cd_px, stream));
}

int srkShrinkageCDBatched(const float* d_E, const float* d_R, int N, int W, int erosion, int B,
                          float coeff_c, int center_row, int L, int A, float threshold,
                          float* d_left_px, float* d_right_px, float* d_cd_px,
                          hipStream_t stream)
{
    if (!d_E || !d_R || !d_left_px || !d_right_px || !d_cd_px) return SRK_ERR_NULL_PTR;
    if (B < 1) return SRK_ERR_BATCH;
    if (W != N - 2 * erosion || W < 2 || erosion < 0 || L < 1 || A < 1) return SRK_ERR_BAD_DIMS;
    if (!srk::cd_shape_ok(W, L, A)) return SRK_ERR_UNSUPPORTED_SIZE;
    return hip_status(srk::cd_measure_shrink(d_E, d_R, N, W, erosion, B, coeff_c, center_row, L, A,
                                             threshold, d_left_px, d_right_px, d_cd_px, stream));
}

}  // extern "C"
