/*
 * srk_oracle.c -- FP64 host oracle for the SRK shrinkage operator.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product (shrinkage_full_tests_amd) never links or calls it.
 *
 * Everything here is double precision and written from the mathematics of the
 * reference operator, citing the reference file:line each function restates:
 *
 *   u_d(p)  = s * sum_{k in [-c,c]^2} g(p - k) * G_d(k)          (true convolution)
 *   h(o)    = u_x(o+e) * f_x(o+e) + u_y(o+e) * f_y(o+e),  o in [0,W)^2
 *   f_x(r,c)= (f(r,c+1) - f(r,c-1))/2,  f_y(r,c) = (f(r+1,c) - f(r-1,c))/2, zero outside
 *   lam_f   = D_x^T P_e[u_x v] + D_y^T P_e[u_y v]                 (stencil transpose)
 *   lam_g(q)= s * sum_d sum_k a_d(q + k) G_d(k),  a_d = P_e[v * f_d]   (correlation)
 *
 * Layout: column-major everywhere, X[c*H + r] (r = row = y, fastest; c = col = x).
 * Kernel layout: K[i*KW + j], i = x (column) offset index, j = y (row) offset index.
 *
 * Parity status of each function is stated in its comment ("pinned by ...").
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* parallel-for over [0, n) with pthreads                                    */
/* ------------------------------------------------------------------------ */
typedef void (*orc_body_fn)(void* ctx, int64_t begin, int64_t end);

static int g_threads = 0;

void orc_set_threads(int n) { g_threads = n; }

int orc_get_threads(void)
{
    if (g_threads > 0) return g_threads;
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

typedef struct {
    orc_body_fn fn;
    void* ctx;
    int64_t begin, end;
} orc_task;

static void* orc_task_run(void* p)
{
    orc_task* t = (orc_task*)p;
    if (t->begin < t->end) t->fn(t->ctx, t->begin, t->end);
    return NULL;
}

static void parallel_for(int64_t n, orc_body_fn fn, void* ctx)
{
    int T = orc_get_threads();
    if (T > 256) T = 256;
    if (n < 2 || T == 1) { fn(ctx, 0, n); return; }
    if (T > n) T = (int)n;
    pthread_t th[256];
    orc_task tasks[256];
    int64_t chunk = (n + T - 1) / T;
    for (int t = 0; t < T; ++t) {
        tasks[t].fn = fn;
        tasks[t].ctx = ctx;
        tasks[t].begin = (int64_t)t * chunk;
        tasks[t].end = tasks[t].begin + chunk < n ? tasks[t].begin + chunk : n;
    }
    for (int t = 1; t < T; ++t) pthread_create(&th[t], NULL, orc_task_run, &tasks[t]);
    orc_task_run(&tasks[0]);
    for (int t = 1; t < T; ++t) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------------ */
/* Modified Bessel K1, FP64.                                                  */
/* Integral representation K_1(x) = int_0^inf exp(-x cosh t) cosh t dt       */
/* (A&S 9.6.24) evaluated with the trapezoid rule, which converges           */
/* geometrically for this entire, even integrand (error ~ exp(-pi^2/h)).     */
/* Replaces the FP32 A&S 9.8.7/9.8.8 polynomials of                          */
/* tan_tanadj/main_tan_tanadj.cu:57-89 and training/main_cd_training.cu:47-62.*/
/* Pinned by: scipy.special.kv(1, x) golden values (tests/golden).           */
/* ------------------------------------------------------------------------ */
double orc_bessel_k1(double x)
{
    if (!(x > 0.0)) return INFINITY;
    const double h = 0.05;
    double sum = 0.5 * exp(-x);
    for (int k = 1; k < 100000; ++k) {
        double t = k * h;
        double ch = cosh(t);
        double term = exp(-x * ch) * ch;
        sum += term;
        if (term < 1e-19 * sum) break;
    }
    return h * sum;
}

/* Radial Blackman window, training/main_cd_training.cu:42-46. */
static double blackman_radial(double r, double Rw)
{
    if (r < 0.0 || r > Rw) return 0.0;
    double t = r / Rw;
    return 0.42 + 0.5 * cos(M_PI * t) + 0.08 * cos(2.0 * M_PI * t);
}

/*
 * SRSRK Green's tensor (Gx, Gy), FP64 restatement of
 * training/main_cd_training.cu:63-77 and tan_tanadj/main_tan_tanadj.cu:93-116:
 *   gamma = 1/sigma, Rw = cw*sigma, coeff = gamma/(2 pi),
 *   x = (i - KW/2)*dx, y = (j - KW/2)*dx, r = |(x,y)|, skip r < 1e-6 dx or r > Rw,
 *   base = coeff*K1(gamma r)*w(r),  Kx[i*KW + j] = base*x/r, Ky = base*y/r.
 * Pinned by: the notebook value Kx(r=1 px, sigma=20) = 0.1583
 * (notebooks/plot_batch_demo.ipynb:86), antisymmetry, Ky = Kx^T, K1 goldens.
 */
void orc_make_srsrk_kernel(int kw, double dx, double sigma, double cw,
                           double* Kx, double* Ky)
{
    const double gamma = 1.0 / sigma;
    const double Rw = cw * sigma;
    const double coeff = gamma / (2.0 * M_PI);
    const double eps = 1e-6 * dx;
    const int cx = kw / 2;
    for (int i = 0; i < kw; ++i)
        for (int j = 0; j < kw; ++j) {
            double x = (i - cx) * dx, y = (j - cx) * dx;
            double r = sqrt(x * x + y * y);
            double kxv = 0.0, kyv = 0.0;
            if (!(r < eps || r > Rw)) {
                double base = coeff * orc_bessel_k1(gamma * r) * blackman_radial(r, Rw);
                kxv = base * (x / r);
                kyv = base * (y / r);
            }
            Kx[(size_t)i * kw + j] = kxv;
            Ky[(size_t)i * kw + j] = kyv;
        }
}

/* ------------------------------------------------------------------------ */
/* Central differences, zero outside [0,N) (srk_forward.cuh:429-445).         */
/* ------------------------------------------------------------------------ */
static inline double fx_at(const double* f, int N, int r, int c)
{
    double right = (c + 1 < N) ? f[(size_t)(c + 1) * N + r] : 0.0;
    double left = (c - 1 >= 0) ? f[(size_t)(c - 1) * N + r] : 0.0;
    return 0.5 * (right - left);
}
static inline double fy_at(const double* f, int N, int r, int c)
{
    double down = (r + 1 < N) ? f[(size_t)c * N + r + 1] : 0.0;
    double up = (r - 1 >= 0) ? f[(size_t)c * N + r - 1] : 0.0;
    return 0.5 * (down - up);
}

/* ------------------------------------------------------------------------ */
/* Direct valid-region convolution (adjoint/main_adjoint.cu:70-90, FP64):     */
/*   u_d(o) = s * sum_k g(o + e - k) G_d(k),  o in [0,W)^2                   */
/* Pinned by: numpy FFT convolution (independent), batch-demo stats.         */
/* ------------------------------------------------------------------------ */
typedef struct {
    const double* g;
    const double* Kx;
    const double* Ky;
    int N, e, kw, W;
    double s;
    double* ux;
    double* uy;
} conv_ctx;

static void conv_body(void* p, int64_t c0, int64_t c1)
{
    conv_ctx* C = (conv_ctx*)p;
    const int N = C->N, e = C->e, kw = C->kw, W = C->W, ck = kw / 2;
    for (int64_t ox = c0; ox < c1; ++ox)
        for (int oy = 0; oy < W; ++oy) {
            int gx = (int)ox + e, gy = oy + e;
            double sx = 0.0, sy = 0.0;
            for (int kx = -ck; kx <= ck; ++kx) {
                int x = gx - kx;
                if (x < 0 || x >= N) continue;
                const double* gcol = C->g + (size_t)x * N;
                const double* kxc = C->Kx + (size_t)(kx + ck) * kw + ck;
                const double* kyc = C->Ky + (size_t)(kx + ck) * kw + ck;
                for (int ky = -ck; ky <= ck; ++ky) {
                    int y = gy - ky;
                    if (y < 0 || y >= N) continue;
                    double gv = gcol[y];
                    sx += gv * kxc[ky];
                    sy += gv * kyc[ky];
                }
            }
            C->ux[(size_t)ox * W + oy] = C->s * sx;
            C->uy[(size_t)ox * W + oy] = C->s * sy;
        }
}

void orc_conv_valid(const double* g, int N, int e, const double* Kx, const double* Ky,
                    int kw, double s, double* ux, double* uy)
{
    conv_ctx C = {g, Kx, Ky, N, e, kw, N - 2 * e, s, ux, uy};
    parallel_for(C.W, conv_body, &C);
}

/* h = u_x f_x + u_y f_y on the valid region (srk_forward.cuh:448-460). */
static void combine(const double* ux, const double* uy, const double* f, int N, int e,
                    double* h)
{
    const int W = N - 2 * e;
    for (int ox = 0; ox < W; ++ox)
        for (int oy = 0; oy < W; ++oy) {
            int gx = ox + e, gy = oy + e;
            size_t o = (size_t)ox * W + oy;
            h[o] = ux[o] * fx_at(f, N, gy, gx) + uy[o] * fy_at(f, N, gy, gx);
        }
}

/*
 * Forward h = s (conv(g,Gx) f_x + conv(g,Gy) f_y) on W x W (srk_forward.cuh:468-501,
 * adjoint/main_adjoint.cu:310-333).  Argument order (f, g) = (I, E).
 * Pinned by: batch-demo known answer (notebooks/plot_batch_demo.ipynb:225-228),
 * numpy FFT convolution, tiling/whole-image agreement of the FFT oracle.
 */
void orc_forward(const double* f, const double* g, int N, int e, const double* Kx,
                 const double* Ky, int kw, double s, double* h)
{
    const int W = N - 2 * e;
    double* ux = (double*)malloc(sizeof(double) * (size_t)W * W);
    double* uy = (double*)malloc(sizeof(double) * (size_t)W * W);
    orc_conv_valid(g, N, e, Kx, Ky, kw, s, ux, uy);
    combine(ux, uy, f, N, e, h);
    free(ux);
    free(uy);
}

/*
 * lam_f = D_x^T(P_e[u_x v]) + D_y^T(P_e[u_y v]) on N x N
 * (adjoint/srk_adjoint.cuh:25-59): with w_d = u_d v (zero outside [0,W)),
 *   lam_f(r,c) = (w_x(r-e, c-e-1) - w_x(r-e, c-e+1))/2 + (w_y(r-e-1, c-e) - w_y(r-e+1, c-e))/2.
 * Pinned by: dot-product test <A_g df, v> = <df, A_g^T v> against orc_forward.
 */
void orc_adjoint_f_from_u(const double* v, const double* ux, const double* uy, int N,
                          int e, double* lam_f)
{
    const int W = N - 2 * e;
    for (int c = 0; c < N; ++c)
        for (int r = 0; r < N; ++r) {
            int ox = c - e, oy = r - e;
            double a = 0.0;
#define WX(rr, cc) (((rr) < 0 || (rr) >= W || (cc) < 0 || (cc) >= W) ? 0.0 \
                    : ux[(size_t)(cc) * W + (rr)] * v[(size_t)(cc) * W + (rr)])
#define WY(rr, cc) (((rr) < 0 || (rr) >= W || (cc) < 0 || (cc) >= W) ? 0.0 \
                    : uy[(size_t)(cc) * W + (rr)] * v[(size_t)(cc) * W + (rr)])
            a = 0.5 * (WX(oy, ox - 1) - WX(oy, ox + 1)) + 0.5 * (WY(oy - 1, ox) - WY(oy + 1, ox));
#undef WX
#undef WY
            lam_f[(size_t)c * N + r] = a;
        }
}

/*
 * True adjoint with respect to g (the reference's CPU brute force
 * adjoint/main_adjoint.cu:206-246, in FP64):
 *   lam_g(q) = s * sum_o v(o) * (Gx(g_o - q) f_x(g_o) + Gy(g_o - q) f_y(g_o)),  g_o = o + e.
 * For odd kernels this is the negative of the reference GPU adjointWrtE
 * (adjoint/srk_adjoint.cuh:301-382), which multiplies by the un-conjugated
 * spectrum; for even kernels the two agree.
 * Pinned by: dot-product test <B_f dg, v> = <dg, B_f^T v> against orc_forward.
 */
typedef struct {
    const double* ax;
    const double* ay;
    const double* Kx;
    const double* Ky;
    int N, kw;
    double s;
    double* lam_g;
} corr_ctx;

static void corr_body(void* p, int64_t c0, int64_t c1)
{
    corr_ctx* C = (corr_ctx*)p;
    const int N = C->N, kw = C->kw, ck = kw / 2;
    for (int64_t qc = c0; qc < c1; ++qc)
        for (int qr = 0; qr < N; ++qr) {
            double acc = 0.0;
            for (int kx = -ck; kx <= ck; ++kx) {
                int pc = (int)qc + kx;
                if (pc < 0 || pc >= N) continue;
                const double* axc = C->ax + (size_t)pc * N;
                const double* ayc = C->ay + (size_t)pc * N;
                const double* kxc = C->Kx + (size_t)(kx + ck) * kw + ck;
                const double* kyc = C->Ky + (size_t)(kx + ck) * kw + ck;
                for (int ky = -ck; ky <= ck; ++ky) {
                    int pr = qr + ky;
                    if (pr < 0 || pr >= N) continue;
                    acc += axc[pr] * kxc[ky] + ayc[pr] * kyc[ky];
                }
            }
            C->lam_g[(size_t)qc * N + qr] = C->s * acc;
        }
}

/* a_d = P_e[v * f_d] on the N x N grid. */
static void make_a(const double* v, const double* f, int N, int e, double* ax, double* ay)
{
    const int W = N - 2 * e;
    memset(ax, 0, sizeof(double) * (size_t)N * N);
    memset(ay, 0, sizeof(double) * (size_t)N * N);
    for (int ox = 0; ox < W; ++ox)
        for (int oy = 0; oy < W; ++oy) {
            int gx = ox + e, gy = oy + e;
            double vv = v[(size_t)ox * W + oy];
            ax[(size_t)gx * N + gy] = vv * fx_at(f, N, gy, gx);
            ay[(size_t)gx * N + gy] = vv * fy_at(f, N, gy, gx);
        }
}

void orc_adjoint_g(const double* v, const double* f, int N, int e, const double* Kx,
                   const double* Ky, int kw, double s, double* lam_g)
{
    double* ax = (double*)malloc(sizeof(double) * (size_t)N * N);
    double* ay = (double*)malloc(sizeof(double) * (size_t)N * N);
    make_a(v, f, N, e, ax, ay);
    corr_ctx C = {ax, ay, Kx, Ky, N, kw, s, lam_g};
    parallel_for(N, corr_body, &C);
    free(ax);
    free(ay);
}

/*
 * adjointSRK with recomputed activations: (lam_f, lam_g) for cotangent v
 * (adjoint/srk_adjoint.cuh:402-425; u recomputed from g instead of stored).
 */
void orc_adjoint(const double* v, const double* f, const double* g, int N, int e,
                 const double* Kx, const double* Ky, int kw, double s, double* lam_f,
                 double* lam_g)
{
    const int W = N - 2 * e;
    double* ux = (double*)malloc(sizeof(double) * (size_t)W * W);
    double* uy = (double*)malloc(sizeof(double) * (size_t)W * W);
    orc_conv_valid(g, N, e, Kx, Ky, kw, s, ux, uy);
    orc_adjoint_f_from_u(v, ux, uy, N, e, lam_f);
    free(ux);
    free(uy);
    orc_adjoint_g(v, f, N, e, Kx, Ky, kw, s, lam_g);
}

/*
 * Tangent dh = u_x(dg) f_x + u_y(dg) f_y + u_x(g) df_x + u_y(g) df_y
 * (tan_tanadj/srk_tangent.cuh:28-52,106-151).
 * Pinned by: bilinearity -- central differences of orc_forward are exact.
 */
void orc_tangent(const double* f, const double* g, const double* df, const double* dg,
                 int N, int e, const double* Kx, const double* Ky, int kw, double s,
                 double* dh)
{
    const int W = N - 2 * e;
    double* t = (double*)malloc(sizeof(double) * (size_t)W * W);
    orc_forward(f, dg, N, e, Kx, Ky, kw, s, dh);
    orc_forward(df, g, N, e, Kx, Ky, kw, s, t);
    for (size_t i = 0; i < (size_t)W * W; ++i) dh[i] += t[i];
    free(t);
}

/*
 * Tangent-adjoint (tan_tanadj/srk_tangent_adjoint.cuh:100-159) with the true
 * adjoint in the g-slot:
 *   dlam_f = A_f(dlam_h; u(g)) + A_f(lam_h; u(dg)),
 *   dlam_g = A_g(dlam_h; f)    + A_g(lam_h; df).
 * Pinned by: bilinearity -- central differences of orc_adjoint are exact.
 */
void orc_tangent_adjoint(const double* f, const double* g, const double* df,
                         const double* dg, const double* lam_h, const double* dlam_h,
                         int N, int e, const double* Kx, const double* Ky, int kw,
                         double s, double* dlam_f, double* dlam_g)
{
    const size_t NN = (size_t)N * N;
    double* tf = (double*)malloc(sizeof(double) * NN);
    double* tg = (double*)malloc(sizeof(double) * NN);
    orc_adjoint(dlam_h, f, g, N, e, Kx, Ky, kw, s, dlam_f, dlam_g);
    orc_adjoint(lam_h, df, dg, N, e, Kx, Ky, kw, s, tf, tg);
    for (size_t i = 0; i < NN; ++i) {
        dlam_f[i] += tf[i];
        dlam_g[i] += tg[i];
    }
    free(tf);
    free(tg);
}

/*
 * Windowed direct forms for large images (restating the same sums as
 * orc_forward / orc_adjoint, evaluated only on a window):
 *   orc_forward_window: h(o) for output pixels o in [ox0, ox0+wx) x [oy0, oy0+wy)
 *     (output = valid-region coordinates), h_out[(ox-ox0)*wy + (oy-oy0)].
 *   orc_adjoint_window: lambda_f, lambda_g at image pixels q in
 *     [qx0, qx0+wx) x [qy0, qy0+wy), out[(qx-qx0)*wy + (qy-qy0)].
 * Pinned by: equality with orc_forward / orc_adjoint on whole small images.
 */
static double u_at(const double* g, int N, const double* K, int kw, int gx, int gy)
{
    const int ck = kw / 2;
    double acc = 0.0;
    for (int kx = -ck; kx <= ck; ++kx) {
        int x = gx - kx;
        if (x < 0 || x >= N) continue;
        for (int ky = -ck; ky <= ck; ++ky) {
            int y = gy - ky;
            if (y < 0 || y >= N) continue;
            acc += g[(size_t)x * N + y] * K[(size_t)(kx + ck) * kw + ky + ck];
        }
    }
    return acc;
}

typedef struct {
    const double *f, *g, *v, *Kx, *Ky;
    int N, e, kw, x0, y0, wx, wy;
    double s;
    double *o0, *o1;
} win_ctx;

static void fwd_win_body(void* p, int64_t a, int64_t b)
{
    win_ctx* C = (win_ctx*)p;
    for (int64_t i = a; i < b; ++i) {
        int ox = C->x0 + (int)(i / C->wy), oy = C->y0 + (int)(i % C->wy);
        int gx = ox + C->e, gy = oy + C->e;
        double ux = C->s * u_at(C->g, C->N, C->Kx, C->kw, gx, gy);
        double uy = C->s * u_at(C->g, C->N, C->Ky, C->kw, gx, gy);
        C->o0[i] = ux * fx_at(C->f, C->N, gy, gx) + uy * fy_at(C->f, C->N, gy, gx);
    }
}

void orc_forward_window(const double* f, const double* g, int N, int e, const double* Kx,
                        const double* Ky, int kw, double s, int ox0, int oy0, int wx, int wy,
                        double* h_out)
{
    win_ctx C = {f, g, NULL, Kx, Ky, N, e, kw, ox0, oy0, wx, wy, s, h_out, NULL};
    parallel_for((int64_t)wx * wy, fwd_win_body, &C);
}

static double cot(const double* v, int N, int e, int c, int r)
{
    const int W = N - 2 * e;
    int oc = c - e, orr = r - e;
    if (oc < 0 || oc >= W || orr < 0 || orr >= W) return 0.0;
    return v[(size_t)oc * W + orr];
}

static void adj_win_body(void* p, int64_t a, int64_t b)
{
    win_ctx* C = (win_ctx*)p;
    const int N = C->N, e = C->e, kw = C->kw, ck = kw / 2;
    for (int64_t i = a; i < b; ++i) {
        int qx = C->x0 + (int)(i / C->wy), qy = C->y0 + (int)(i % C->wy);
        /* lambda_f: stencil transpose of w_d = u_d v */
        double wxm = cot(C->v, N, e, qx - 1, qy), wxp = cot(C->v, N, e, qx + 1, qy);
        double wym = cot(C->v, N, e, qx, qy - 1), wyp = cot(C->v, N, e, qx, qy + 1);
        double lf = 0.0;
        if (wxm != 0.0) lf += 0.5 * C->s * u_at(C->g, N, C->Kx, kw, qx - 1, qy) * wxm;
        if (wxp != 0.0) lf -= 0.5 * C->s * u_at(C->g, N, C->Kx, kw, qx + 1, qy) * wxp;
        if (wym != 0.0) lf += 0.5 * C->s * u_at(C->g, N, C->Ky, kw, qx, qy - 1) * wym;
        if (wyp != 0.0) lf -= 0.5 * C->s * u_at(C->g, N, C->Ky, kw, qx, qy + 1) * wyp;
        C->o0[i] = lf;
        /* lambda_g: correlation of a_d = v f_d with G_d */
        double lg = 0.0;
        for (int kx = -ck; kx <= ck; ++kx) {
            int pc = qx + kx;
            if (pc < e || pc >= N - e) continue;
            for (int ky = -ck; ky <= ck; ++ky) {
                int pr = qy + ky;
                if (pr < e || pr >= N - e) continue;
                double vv = cot(C->v, N, e, pc, pr);
                lg += vv * (fx_at(C->f, N, pr, pc) * C->Kx[(size_t)(kx + ck) * kw + ky + ck] +
                            fy_at(C->f, N, pr, pc) * C->Ky[(size_t)(kx + ck) * kw + ky + ck]);
            }
        }
        C->o1[i] = C->s * lg;
    }
}

void orc_adjoint_window(const double* v, const double* f, const double* g, int N, int e,
                        const double* Kx, const double* Ky, int kw, double s, int qx0, int qy0,
                        int wx, int wy, double* lam_f, double* lam_g)
{
    win_ctx C = {f, g, v, Kx, Ky, N, e, kw, qx0, qy0, wx, wy, s, lam_f, lam_g};
    parallel_for((int64_t)wx * wy, adj_win_body, &C);
}

/* ------------------------------------------------------------------------ */
/* FP64 FFT path (the timed CPU baseline).  Radix-2 iterative complex FFT,    */
/* 2-D by lines, multithreaded over lines.  Whole-image circular convolution  */
/* at M >= N (power of two) is exact on the regions used because erosion >=   */
/* (KW-1)/2 (see the notes in orc_forward_fft / orc_adjoint_fft).             */
/* Pinned by: agreement with the direct functions above.                      */
/* ------------------------------------------------------------------------ */
static void fft1(double* re, double* im, int n, int sign)
{
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        double ang = sign * 2.0 * M_PI / len;
        for (int k = 0; k < len / 2; ++k) {
            double wr = cos(ang * k), wi = sin(ang * k);
            for (int i = 0; i < n; i += len) {
                int a = i + k, b = i + k + len / 2;
                double xr = re[b] * wr - im[b] * wi;
                double xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr; im[b] = im[a] - xi;
                re[a] += xr; im[a] += xi;
            }
        }
    }
}

typedef struct {
    double* re;
    double* im;
    int M;
    int sign;
    int axis; /* 0: along rows index (contiguous, per column); 1: along columns */
} fft2_ctx;

static void fft2_body(void* p, int64_t l0, int64_t l1)
{
    fft2_ctx* C = (fft2_ctx*)p;
    const int M = C->M;
    double* tr = (double*)malloc(sizeof(double) * M);
    double* ti = (double*)malloc(sizeof(double) * M);
    for (int64_t l = l0; l < l1; ++l) {
        if (C->axis == 0) {
            fft1(C->re + (size_t)l * M, C->im + (size_t)l * M, M, C->sign);
        } else {
            for (int k = 0; k < M; ++k) {
                tr[k] = C->re[(size_t)k * M + l];
                ti[k] = C->im[(size_t)k * M + l];
            }
            fft1(tr, ti, M, C->sign);
            for (int k = 0; k < M; ++k) {
                C->re[(size_t)k * M + l] = tr[k];
                C->im[(size_t)k * M + l] = ti[k];
            }
        }
    }
    free(tr);
    free(ti);
}

static void fft2(double* re, double* im, int M, int sign)
{
    fft2_ctx a = {re, im, M, sign, 0};
    parallel_for(M, fft2_body, &a);
    fft2_ctx b = {re, im, M, sign, 1};
    parallel_for(M, fft2_body, &b);
}

static int next_pow2(int n)
{
    int m = 1;
    while (m < n) m <<= 1;
    return m;
}

/* kernel placed at circular offsets: K(k) -> index (k mod M) */
static void kernel_spectrum(const double* K, int kw, int M, double* re, double* im)
{
    const int ck = kw / 2;
    memset(re, 0, sizeof(double) * (size_t)M * M);
    memset(im, 0, sizeof(double) * (size_t)M * M);
    for (int i = 0; i < kw; ++i)
        for (int j = 0; j < kw; ++j) {
            int x = ((i - ck) % M + M) % M, y = ((j - ck) % M + M) % M;
            re[(size_t)x * M + y] = K[(size_t)i * kw + j];
        }
    fft2(re, im, M, -1);
}

/*
 * FFT form of orc_forward: circular convolution at M = next_pow2(N) >= N.
 * For outputs p in [e, N-e) the taps p - k stay inside [0, N) (e >= c), so no
 * wrap-around contributes.
 */
void orc_forward_fft(const double* f, const double* g, int N, int e, const double* Kx,
                     const double* Ky, int kw, double s, double* h)
{
    const int M = next_pow2(N), W = N - 2 * e;
    const size_t MM = (size_t)M * M;
    double* gr = (double*)calloc(MM, sizeof(double));
    double* gi = (double*)calloc(MM, sizeof(double));
    double* kr = (double*)malloc(sizeof(double) * MM);
    double* ki = (double*)malloc(sizeof(double) * MM);
    double* ur = (double*)malloc(sizeof(double) * MM);
    double* ui = (double*)malloc(sizeof(double) * MM);
    double* ux = (double*)malloc(sizeof(double) * (size_t)W * W);
    double* uy = (double*)malloc(sizeof(double) * (size_t)W * W);
    for (int c = 0; c < N; ++c) memcpy(gr + (size_t)c * M, g + (size_t)c * N, sizeof(double) * N);
    fft2(gr, gi, M, -1);
    const double sc = s / ((double)M * M);
    for (int d = 0; d < 2; ++d) {
        kernel_spectrum(d == 0 ? Kx : Ky, kw, M, kr, ki);
        for (size_t i = 0; i < MM; ++i) {
            ur[i] = sc * (gr[i] * kr[i] - gi[i] * ki[i]);
            ui[i] = sc * (gr[i] * ki[i] + gi[i] * kr[i]);
        }
        fft2(ur, ui, M, +1);
        double* u = d == 0 ? ux : uy;
        for (int ox = 0; ox < W; ++ox)
            for (int oy = 0; oy < W; ++oy) u[(size_t)ox * W + oy] = ur[(size_t)(ox + e) * M + oy + e];
    }
    combine(ux, uy, f, N, e, h);
    free(gr); free(gi); free(kr); free(ki); free(ur); free(ui); free(ux); free(uy);
}

/*
 * FFT form of orc_adjoint.  lam_g is a circular correlation at M >= N with
 * conj(G^).  Exact on [0,N): a_d vanishes outside [e, N-e), and every wrapped
 * tap index falls into that zero band because c <= e.
 */
void orc_adjoint_fft(const double* v, const double* f, const double* g, int N, int e,
                     const double* Kx, const double* Ky, int kw, double s, double* lam_f,
                     double* lam_g)
{
    const int M = next_pow2(N), W = N - 2 * e;
    const size_t MM = (size_t)M * M, NN = (size_t)N * N;
    double* ux = (double*)malloc(sizeof(double) * (size_t)W * W);
    double* uy = (double*)malloc(sizeof(double) * (size_t)W * W);
    double* gr = (double*)calloc(MM, sizeof(double));
    double* gi = (double*)calloc(MM, sizeof(double));
    double* kxr = (double*)malloc(sizeof(double) * MM);
    double* kxi = (double*)malloc(sizeof(double) * MM);
    double* kyr = (double*)malloc(sizeof(double) * MM);
    double* kyi = (double*)malloc(sizeof(double) * MM);
    double* ur = (double*)malloc(sizeof(double) * MM);
    double* ui = (double*)malloc(sizeof(double) * MM);
    double* ax = (double*)malloc(sizeof(double) * NN);
    double* ay = (double*)malloc(sizeof(double) * NN);
    const double sc = s / ((double)M * M);
    kernel_spectrum(Kx, kw, M, kxr, kxi);
    kernel_spectrum(Ky, kw, M, kyr, kyi);
    /* recompute u */
    for (int c = 0; c < N; ++c) memcpy(gr + (size_t)c * M, g + (size_t)c * N, sizeof(double) * N);
    fft2(gr, gi, M, -1);
    for (int d = 0; d < 2; ++d) {
        const double* kr = d == 0 ? kxr : kyr;
        const double* ki = d == 0 ? kxi : kyi;
        for (size_t i = 0; i < MM; ++i) {
            ur[i] = sc * (gr[i] * kr[i] - gi[i] * ki[i]);
            ui[i] = sc * (gr[i] * ki[i] + gi[i] * kr[i]);
        }
        fft2(ur, ui, M, +1);
        double* u = d == 0 ? ux : uy;
        for (int ox = 0; ox < W; ++ox)
            for (int oy = 0; oy < W; ++oy) u[(size_t)ox * W + oy] = ur[(size_t)(ox + e) * M + oy + e];
    }
    orc_adjoint_f_from_u(v, ux, uy, N, e, lam_f);
    /* lam_g: pack a_x + i a_y, one complex transform, unpack with Hermitian symmetry */
    make_a(v, f, N, e, ax, ay);
    memset(gr, 0, sizeof(double) * MM);
    memset(gi, 0, sizeof(double) * MM);
    for (int c = 0; c < N; ++c)
        for (int r = 0; r < N; ++r) {
            gr[(size_t)c * M + r] = ax[(size_t)c * N + r];
            gi[(size_t)c * M + r] = ay[(size_t)c * N + r];
        }
    fft2(gr, gi, M, -1);
    for (int x = 0; x < M; ++x)
        for (int y = 0; y < M; ++y) {
            size_t i = (size_t)x * M + y;
            size_t j = (size_t)((M - x) % M) * M + (M - y) % M;
            /* A_x = (Z(k) + conj Z(-k))/2, A_y = (Z(k) - conj Z(-k))/(2i) */
            double axr = 0.5 * (gr[i] + gr[j]), axi = 0.5 * (gi[i] - gi[j]);
            double ayr = 0.5 * (gi[i] + gi[j]), ayi = -0.5 * (gr[i] - gr[j]);
            /* L = conj(Kx) A_x + conj(Ky) A_y */
            ur[i] = sc * (kxr[i] * axr + kxi[i] * axi + kyr[i] * ayr + kyi[i] * ayi);
            ui[i] = sc * (kxr[i] * axi - kxi[i] * axr + kyr[i] * ayi - kyi[i] * ayr);
        }
    fft2(ur, ui, M, +1);
    for (int c = 0; c < N; ++c)
        for (int r = 0; r < N; ++r) lam_g[(size_t)c * N + r] = ur[(size_t)c * M + r];
    free(ux); free(uy); free(gr); free(gi); free(kxr); free(kxi); free(kyr); free(kyi);
    free(ur); free(ui); free(ax); free(ay);
}

/* Batched FFT-form forward + adjoint over B images (the CPU baseline unit). */
void orc_fwd_adj_fft_batch(const double* f, const double* g, const double* v, int B, int N,
                           int e, const double* Kx, const double* Ky, int kw, double s,
                           double* h, double* lam_f, double* lam_g)
{
    const int W = N - 2 * e;
    const size_t NN = (size_t)N * N, WW = (size_t)W * W;
    for (int b = 0; b < B; ++b) {
        orc_forward_fft(f + b * NN, g + b * NN, N, e, Kx, Ky, kw, s, h + b * WW);
        orc_adjoint_fft(v + b * WW, f + b * NN, g + b * NN, N, e, Kx, Ky, kw, s,
                        lam_f + b * NN, lam_g + b * NN);
    }
}
