/*
 * CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Real-to-real transforms with FFTW 3.3's unnormalised definitions (the
 * reference links FFTW 3.3.4 / 3.3.8, see runAdv:28 and runQuasi:27; FFTW is
 * not vendored in the reference, so its published definitions are restated):
 *
 *   REDFT10 (DCT-II):  Y_k = 2 sum_n x_n cos(pi k (2n+1) / 2N)
 *   REDFT01 (DCT-III): Y_k = X_0 + 2 sum_{n>=1} X_n cos(pi n (2k+1) / 2N)
 *   RODFT10 (DST-II):  Y_k = 2 sum_n x_n sin(pi (k+1)(2n+1) / 2N)
 *   RODFT01 (DST-III): Y_k = (-1)^k X_{N-1} + 2 sum_{n<=N-2} X_n sin(pi (n+1)(2k+1) / 2N)
 *
 * Call sites in the reference: no-adv/cosNoAdvTrackSing.cpp:342-375,
 * adv/cosAdvTrackMult.cpp:419-467, quasi/quasi.cpp:519-565.
 *
 * Two implementations: orc_r2r1d_naive() evaluates the sums directly with an
 * exact integer argument reduction (O(N^2)); orc_r2r1d() uses Makhoul's
 * even/odd reordering (IEEE TASSP 28(1), 1980) around a recursive mixed-radix
 * complex FFT (O(N log N)).  Both are pinned against scipy.fft fixtures and
 * against each other by the CPU tests.  Nothing here is shipped or called by
 * the product library.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------- naive -- */

/* cos(pi*m/(2N)) for m in [0,4N): exact periodic reduction of the argument. */
static double cospi_2n(long m, int n)
{
    long p = 4L * n;
    m %= p;
    if (m < 0) m += p;
    return cos(M_PI * (double)m / (2.0 * n));
}
static double sinpi_2n(long m, int n)
{
    long p = 4L * n;
    m %= p;
    if (m < 0) m += p;
    return sin(M_PI * (double)m / (2.0 * n));
}

void orc_r2r1d_naive(const double *x, double *y, int n, int kind)
{
    for (int k = 0; k < n; ++k) {
        double acc = 0.0;
        switch (kind) {
        case ORC_REDFT10:
            for (int j = 0; j < n; ++j) acc += x[j] * cospi_2n((long)k * (2 * j + 1), n);
            acc *= 2.0;
            break;
        case ORC_REDFT01:
            for (int j = 1; j < n; ++j) acc += x[j] * cospi_2n((long)j * (2 * k + 1), n);
            acc = x[0] + 2.0 * acc;
            break;
        case ORC_RODFT10:
            for (int j = 0; j < n; ++j) acc += x[j] * sinpi_2n((long)(k + 1) * (2 * j + 1), n);
            acc *= 2.0;
            break;
        case ORC_RODFT01:
            for (int j = 0; j < n - 1; ++j) acc += x[j] * sinpi_2n((long)(j + 1) * (2 * k + 1), n);
            acc = ((k & 1) ? -x[n - 1] : x[n - 1]) + 2.0 * acc;
            break;
        default:
            break;
        }
        y[k] = acc;
    }
}

/* ------------------------------------------------------------------ fast -- */

typedef struct { double re, im; } cplx;

static inline cplx cmul(cplx a, cplx b)
{
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}

/* Per-length tables: w[j] = e^{-2 pi i j/N}, h[k] = e^{-i pi k/2N}. */
typedef struct {
    int n;
    cplx *w;
    cplx *h;
    cplx *v, *tmp, *scr; /* work buffers (one plan per thread) */
    double *t;
} r2r_plan;

static r2r_plan *plan_make(int n)
{
    r2r_plan *p = (r2r_plan *)calloc(1, sizeof(r2r_plan));
    p->n = n;
    p->w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    p->h = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int j = 0; j < n; ++j) {
        p->w[j].re = cos(2.0 * M_PI * j / n);
        p->w[j].im = -sin(2.0 * M_PI * j / n);
        p->h[j].re = cos(M_PI * j / (2.0 * n));
        p->h[j].im = -sin(M_PI * j / (2.0 * n));
    }
    p->v = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    p->tmp = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    p->scr = (cplx *)malloc(sizeof(cplx) * (size_t)n * 4);
    p->t = (double *)malloc(sizeof(double) * (size_t)n);
    return p;
}

static void plan_free(r2r_plan *p)
{
    free(p->w); free(p->h); free(p->v); free(p->tmp); free(p->scr); free(p->t);
    free(p);
}

/* twiddle e^{sign 2 pi i e/m} from the length-N table (m divides N) */
static inline cplx tw(const r2r_plan *p, long e, int m, int sign)
{
    e %= m;
    cplx w = p->w[e * (p->n / m)];
    if (sign > 0) w.im = -w.im;
    return w;
}

/* Recursive mixed-radix decimation-in-time DFT:
 * out[k] = sum_j in[j*stride] e^{sign 2 pi i jk/m}.  Radices 4, 2, 3, 5, and a
 * direct DFT for any other prime factor. */
static void fft_rec(const r2r_plan *p, const cplx *in, long stride, cplx *out, int m, int sign,
                    cplx *scratch)
{
    if (m == 1) { out[0] = in[0]; return; }
    int r;
    if (m % 4 == 0) r = 4;
    else if (m % 2 == 0) r = 2;
    else if (m % 3 == 0) r = 3;
    else if (m % 5 == 0) r = 5;
    else r = m;
    int q = m / r;
    if (r == m) {
        for (int k = 0; k < m; ++k) {
            cplx acc = {0.0, 0.0};
            for (int j = 0; j < m; ++j) {
                cplx t = cmul(in[j * stride], tw(p, (long)j * k, m, sign));
                acc.re += t.re;
                acc.im += t.im;
            }
            scratch[k] = acc;
        }
        memcpy(out, scratch, sizeof(cplx) * (size_t)m);
        return;
    }
    for (int s = 0; s < r; ++s)
        fft_rec(p, in + s * stride, stride * r, out + s * q, q, sign, scratch + m);
    for (int k = 0; k < q; ++k) {
        cplx ys[5];
        for (int s = 0; s < r; ++s) ys[s] = cmul(out[s * q + k], tw(p, (long)s * k, m, sign));
        for (int t = 0; t < r; ++t) {
            cplx acc = {0.0, 0.0};
            for (int s = 0; s < r; ++s) {
                cplx u = cmul(ys[s], tw(p, (long)s * t, r, sign));
                acc.re += u.re;
                acc.im += u.im;
            }
            scratch[k + t * q] = acc;
        }
    }
    memcpy(out, scratch, sizeof(cplx) * (size_t)m);
}

static void fft(r2r_plan *p, cplx *a, int sign)
{
    memcpy(p->tmp, a, sizeof(cplx) * (size_t)p->n);
    fft_rec(p, p->tmp, 1, a, p->n, sign, p->scr);
}

/* DCT-II by Makhoul: v_j = x_{2j}, v_{N-1-j} = x_{2j+1}; V = DFT_N(v);
 * Y_k = 2 Re(e^{-i pi k/2N} V_k).  A full N-point complex DFT is used here
 * (no half-length packing), which keeps this path independent of the
 * device kernels' packed form. */
static void dct2_fast(r2r_plan *p, const double *x, double *y)
{
    int n = p->n;
    cplx *v = p->v;
    for (int j = 0; j < (n + 1) / 2; ++j) { v[j].re = x[2 * j]; v[j].im = 0.0; }
    for (int j = 0; j < n / 2; ++j) { v[n - 1 - j].re = x[2 * j + 1]; v[n - 1 - j].im = 0.0; }
    fft(p, v, -1);
    for (int k = 0; k < n; ++k) y[k] = 2.0 * cmul(p->h[k], v[k]).re;
}

/* DCT-III as the inverse of the above: V_k = e^{+i pi k/2N} (X_k - i X_{N-k}) (X_N = 0),
 * v_j = sum_k V_k e^{+2 pi i jk/N} (real), then undo the reordering.  This is
 * REDFT01 exactly because REDFT01 o REDFT10 = 2N I. */
static void dct3_fast(r2r_plan *p, const double *x, double *y)
{
    int n = p->n;
    cplx *v = p->v;
    for (int k = 0; k < n; ++k) {
        cplx w = p->h[k];
        w.im = -w.im;
        cplx z = {x[k], k == 0 ? 0.0 : -x[n - k]};
        v[k] = cmul(w, z);
    }
    fft(p, v, +1);
    for (int j = 0; j < (n + 1) / 2; ++j) y[2 * j] = v[j].re;
    for (int j = 0; j < n / 2; ++j) y[2 * j + 1] = v[n - 1 - j].re;
}

static void r2r_plan_exec(r2r_plan *p, const double *x, double *y, int kind)
{
    int n = p->n;
    double *t = p->t;
    switch (kind) {
    case ORC_REDFT10:
        dct2_fast(p, x, y);
        break;
    case ORC_REDFT01:
        dct3_fast(p, x, y);
        break;
    case ORC_RODFT10:
        /* RODFT10(x)_k = REDFT10((-1)^j x_j)_{N-1-k} */
        for (int j = 0; j < n; ++j) t[j] = (j & 1) ? -x[j] : x[j];
        dct2_fast(p, t, t);
        for (int k = 0; k < n; ++k) y[k] = t[n - 1 - k];
        break;
    case ORC_RODFT01:
        /* RODFT01(X)_k = (-1)^k REDFT01(X reversed)_k */
        for (int j = 0; j < n; ++j) t[j] = x[n - 1 - j];
        dct3_fast(p, t, t);
        for (int k = 0; k < n; ++k) y[k] = (k & 1) ? -t[k] : t[k];
        break;
    default:
        break;
    }
}

void orc_r2r1d(const double *x, double *y, int n, int kind)
{
    r2r_plan *p = plan_make(n);
    double *tmp = (double *)malloc(sizeof(double) * (size_t)n);
    r2r_plan_exec(p, x, tmp, kind);
    memcpy(y, tmp, sizeof(double) * (size_t)n);
    free(tmp);
    plan_free(p);
}

/* ------------------------------------------------------------------- 3D -- */

/* One axis of a separable 3D transform on a row-major (i,j,k) array, k fastest. */
static void axis_pass(double *a, int nx, int ny, int nz, int axis, int kind)
{
    long stride, nlines;
    int n;
    if (axis == 2) { n = nz; stride = 1; nlines = (long)nx * ny; }
    else if (axis == 1) { n = ny; stride = nz; nlines = (long)nx * nz; }
    else { n = nx; stride = (long)ny * nz; nlines = (long)ny * nz; }

#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * (size_t)n);
        double *out = (double *)malloc(sizeof(double) * (size_t)n);
        r2r_plan *pl = plan_make(n);
#pragma omp for schedule(static)
        for (long l = 0; l < nlines; ++l) {
            long base;
            if (axis == 2) base = l * nz;
            else if (axis == 1) base = (l / nz) * (long)ny * nz + (l % nz);
            else base = l;
            for (int p = 0; p < n; ++p) buf[p] = a[base + p * stride];
            r2r_plan_exec(pl, buf, out, kind);
            for (int p = 0; p < n; ++p) a[base + p * stride] = out[p];
        }
        free(buf);
        free(out);
        plan_free(pl);
    }
}

void orc_r2r3d(const double *in, double *out, int nx, int ny, int nz, int kx, int ky, int kz)
{
    if (out != in) memcpy(out, in, sizeof(double) * (size_t)nx * ny * nz);
    axis_pass(out, nx, ny, nz, 2, kz);
    axis_pass(out, nx, ny, nz, 1, ky);
    axis_pass(out, nx, ny, nz, 0, kx);
}

/* Multiplies the plane `at` (0 or n-1) of one axis by w. */
static void weight_plane(double *a, int nx, int ny, int nz, int axis, int at, double w)
{
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                int c = axis == 0 ? i : axis == 1 ? j : k;
                if (c == at) a[((long)i * ny + j) * nz + k] *= w;
            }
}

/* Transpose of the separable transform with kinds (kx, ky, kz), axis by
 * axis: REDFT10^T = REDFT01 diag(2,1,..,1), REDFT01^T = diag(1/2,1,..,1)
 * REDFT10, RODFT10^T = RODFT01 diag(1,..,1,2), RODFT01^T = diag(1,..,1,1/2)
 * RODFT10 (the unnormalised FFTW kinds). */
void orc_r2r3d_T(const double *in, double *out, int nx, int ny, int nz, int kx, int ky, int kz)
{
    if (out != in) memcpy(out, in, sizeof(double) * (size_t)nx * ny * nz);
    const int kinds[3] = {kx, ky, kz}, len[3] = {nx, ny, nz};
    for (int axis = 0; axis < 3; ++axis) {
        const int last = len[axis] - 1;
        switch (kinds[axis]) {
        case ORC_REDFT10:
            weight_plane(out, nx, ny, nz, axis, 0, 2.0);
            axis_pass(out, nx, ny, nz, axis, ORC_REDFT01);
            break;
        case ORC_REDFT01:
            axis_pass(out, nx, ny, nz, axis, ORC_REDFT10);
            weight_plane(out, nx, ny, nz, axis, 0, 0.5);
            break;
        case ORC_RODFT10:
            weight_plane(out, nx, ny, nz, axis, last, 2.0);
            axis_pass(out, nx, ny, nz, axis, ORC_RODFT01);
            break;
        default:
            axis_pass(out, nx, ny, nz, axis, ORC_RODFT10);
            weight_plane(out, nx, ny, nz, axis, last, 0.5);
            break;
        }
    }
}
