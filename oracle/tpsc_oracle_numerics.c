/*
 * tpsc_oracle_numerics.c -- numerical kernels of the oracle (test infrastructure only).
 *
 * Restates generique.cpp of dbergeron1/TPSC_conductivity (root finder, clamped spline,
 * quadratic roots, Pade) and the transform semantics of the third-party FFTW3 calls the
 * reference makes (REDFT00, RODFT00, complex DFT; FFTW 3.x public definitions, version
 * unpinned by the reference: src/CMakeLists.txt:11).  LAPACK dgbsv_ (also unpinned) is
 * restated as the unblocked banded LU with partial pivoting of LAPACK's dgbtf2 followed
 * by the dgbtrs/dtbsv substitutions -- the path dgbsv takes for KL=3, KU=2.
 */
#include "tpsc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.14159265358979323846

static void *orc_xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "tpsc_oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}

/* ------------------------------------------------------------------------------------
 * Transforms.  FFTW definitions (unnormalised):
 *   REDFT00 (n>=2): Y_k = X_0 + (-1)^k X_{n-1} + 2 sum_{j=1}^{n-2} X_j cos(pi j k/(n-1))
 *   RODFT00       : Y_k = 2 sum_{j=0}^{n-1} X_j sin(pi (j+1)(k+1)/(n+1))
 *   dft sign s    : Y_k = sum_j X_j exp(s * 2 pi i j k / n)   (s=+1 FFTW_BACKWARD)
 * 2D transforms are applied by dense transform matrices (one per axis), which is the
 * definition itself; the matrices are cached per (n, kind).
 * ---------------------------------------------------------------------------------- */
typedef struct {
    int n, kind;
    double *M;   /* M[k*n + j]  (output k, input j) */
    double *MT;  /* MT[j*n + k]                      */
} mat_entry;

#define MAX_MAT 64
static mat_entry g_mat[MAX_MAT];
static int g_nmat = 0;

static const mat_entry *get_mat(int n, int kind)
{
    const mat_entry *res = NULL;
#pragma omp critical(orc_mat_cache)
    {
        for (int i = 0; i < g_nmat; i++)
            if (g_mat[i].n == n && g_mat[i].kind == kind) res = &g_mat[i];
        if (!res) {
            if (g_nmat == MAX_MAT) {
                fprintf(stderr, "tpsc_oracle: transform cache full\n");
                abort();
            }
            mat_entry *e = &g_mat[g_nmat];
            e->n = n;
            e->kind = kind;
            e->M = (double *)orc_xmalloc(sizeof(double) * (size_t)n * n);
            e->MT = (double *)orc_xmalloc(sizeof(double) * (size_t)n * n);
            for (int k = 0; k < n; k++)
                for (int j = 0; j < n; j++) {
                    double v;
                    if (kind == 0) {
                        long per = 2L * (n - 1);
                        long r = ((long)j * k) % per;
                        double w = (j == 0 || j == n - 1) ? 1.0 : 2.0;
                        v = w * cos(PI * (double)r / (double)(n - 1));
                    } else {
                        long per = 2L * (n + 1);
                        long r = ((long)(j + 1) * (k + 1)) % per;
                        v = 2.0 * sin(PI * (double)r / (double)(n + 1));
                    }
                    e->M[(size_t)k * n + j] = v;
                    e->MT[(size_t)j * n + k] = v;
                }
            g_nmat++;
            res = e;
        }
    }
    return res;
}

void orc_r2r_2d(double *a, int n0, int n1, int kind0, int kind1)
{
    const mat_entry *m1 = get_mat(n1, kind1);
    const mat_entry *m0 = get_mat(n0, kind0);
    double *tmp = (double *)orc_xmalloc(sizeof(double) * (size_t)n0 * n1);
    /* axis 1 (fast index): tmp[i0][k] = sum_j a[i0][j] M1[k][j] */
    for (int i0 = 0; i0 < n0; i0++) {
        double *t = tmp + (size_t)i0 * n1;
        const double *src = a + (size_t)i0 * n1;
        for (int k = 0; k < n1; k++) t[k] = 0.0;
        for (int j = 0; j < n1; j++) {
            const double x = src[j];
            const double *row = m1->MT + (size_t)j * n1;
            for (int k = 0; k < n1; k++) t[k] += x * row[k];
        }
    }
    /* axis 0 (slow index): a[k][i1] = sum_j M0[k][j] tmp[j][i1] */
    for (int k = 0; k < n0; k++) {
        double *o = a + (size_t)k * n1;
        for (int i1 = 0; i1 < n1; i1++) o[i1] = 0.0;
        for (int j = 0; j < n0; j++) {
            const double c = m0->M[(size_t)k * n0 + j];
            const double *t = tmp + (size_t)j * n1;
            for (int i1 = 0; i1 < n1; i1++) o[i1] += c * t[i1];
        }
    }
    free(tmp);
}

/* twiddle tables w[k] = exp(2 pi i k / n), cached per n */
typedef struct {
    int n;
    double complex *w;
} tw_entry;
#define MAX_TW 64
static tw_entry g_tw[MAX_TW];
static int g_ntw = 0;

static const double complex *get_tw(int n)
{
    const double complex *res = NULL;
#pragma omp critical(orc_tw_cache)
    {
        for (int i = 0; i < g_ntw; i++)
            if (g_tw[i].n == n) res = g_tw[i].w;
        if (!res) {
            if (g_ntw == MAX_TW) {
                fprintf(stderr, "tpsc_oracle: twiddle cache full\n");
                abort();
            }
            double complex *w = (double complex *)orc_xmalloc(sizeof(double complex) * (size_t)n);
            for (int k = 0; k < n; k++) {
                double ang = 2.0 * PI * (double)k / (double)n;
                w[k] = cos(ang) + I * sin(ang);
            }
            g_tw[g_ntw].n = n;
            g_tw[g_ntw].w = w;
            g_ntw++;
            res = w;
        }
    }
    return res;
}

void orc_dft(double complex *a, int n, int sign)
{
    if (n <= 1) return;
    const double complex *w = get_tw(n);
    if (n & (n - 1)) { /* not a power of two: direct definition */
        double complex *out = (double complex *)orc_xmalloc(sizeof(double complex) * (size_t)n);
        for (int k = 0; k < n; k++) {
            double complex s = 0;
            for (int j = 0; j < n; j++) {
                double complex t = w[((long)j * k) % n];
                s += a[j] * (sign > 0 ? t : conj(t));
            }
            out[k] = s;
        }
        memcpy(a, out, sizeof(double complex) * (size_t)n);
        free(out);
        return;
    }
    /* iterative radix-2 decimation in time */
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double complex t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len)
            for (int j = 0; j < half; j++) {
                double complex t = w[(long)j * step];
                if (sign < 0) t = conj(t);
                double complex u = a[i + j], v = a[i + j + half] * t;
                a[i + j] = u + v;
                a[i + j + half] = u - v;
            }
    }
}

void orc_redft00_1d(double *a, int n)
{
    int m = n - 1;
    if (m >= 2 && (m & (m - 1)) == 0) {
        /* even extension of length 2m; its DFT is real and equals REDFT00 */
        int L = 2 * m;
        double complex *z = (double complex *)orc_xmalloc(sizeof(double complex) * (size_t)L);
        for (int j = 0; j <= m; j++) z[j] = a[j];
        for (int j = 1; j < m; j++) z[L - j] = a[j];
        orc_dft(z, L, -1);
        for (int k = 0; k <= m; k++) a[k] = creal(z[k]);
        free(z);
    } else {
        const mat_entry *me = get_mat(n, 0);
        double *t = (double *)orc_xmalloc(sizeof(double) * (size_t)n);
        for (int k = 0; k < n; k++) {
            double s = 0;
            for (int j = 0; j < n; j++) s += me->M[(size_t)k * n + j] * a[j];
            t[k] = s;
        }
        memcpy(a, t, sizeof(double) * (size_t)n);
        free(t);
    }
}

/* ------------------------------------------------------------------------------------
 * Clamped cubic spline, generique.cpp:214-322.
 * Unknown vector u = (a0, b0, a1, b1, c1, ..., a_j, b_j, c_j, ...), u[3j-1]=a_j,
 * u[3j]=b_j, u[3j+1]=c_j for j>=1; c0 = coeffs[0] (left slope) and d_j = V[j] known.
 * Rows: value continuity h^3 a_j + h^2 b_j + h c_j = V_{j+1}-V_j; second-derivative
 * continuity; first-derivative continuity; last row: right slope.  Band KL=3, KU=2.
 * ---------------------------------------------------------------------------------- */
#define SKL 3
#define SKU 2
#define SW (2 * SKL + SKU + 1)

static inline double *band_at(double *B, long i, long c) { return &B[i * SW + (c - i + SKL)]; }

void orc_spline_coeffs_rel_ws(const double *x0, const double *V, int N0, double *coeffs,
                              double *B, double *r)
{
    const int NS = N0 - 1;
    const int N = 3 * NS - 1;
    const double dV1 = coeffs[0], dVN0 = coeffs[1];
    memset(B, 0, sizeof(double) * (size_t)N * SW);
    memset(r, 0, sizeof(double) * (size_t)N);

    double h = x0[1] - x0[0];
    *band_at(B, 0, 0) = h * h * h;
    *band_at(B, 1, 0) = -6.0 * h;
    *band_at(B, 2, 0) = -3.0 * h * h;
    *band_at(B, 0, 1) = h * h;
    *band_at(B, 1, 1) = -2.0;
    *band_at(B, 2, 1) = -2.0 * h;
    r[0] = V[1] - V[0] - dV1 * h;
    r[2] = dV1;
    for (int j = 1; j < NS - 1; j++) {
        h = x0[j + 1] - x0[j];
        *band_at(B, 3 * j, 3 * j - 1) = h * h * h;
        *band_at(B, 3 * j + 1, 3 * j - 1) = -6.0 * h;
        *band_at(B, 3 * j + 2, 3 * j - 1) = -3.0 * h * h;
        *band_at(B, 3 * j - 2, 3 * j) = 2.0;
        *band_at(B, 3 * j, 3 * j) = h * h;
        *band_at(B, 3 * j + 1, 3 * j) = -2.0;
        *band_at(B, 3 * j + 2, 3 * j) = -2.0 * h;
        *band_at(B, 3 * j - 1, 3 * j + 1) = 1.0;
        *band_at(B, 3 * j, 3 * j + 1) = h;
        *band_at(B, 3 * j + 2, 3 * j + 1) = -1.0;
        r[3 * j] = V[j + 1] - V[j];
    }
    {
        const int j = NS - 1;
        h = x0[j + 1] - x0[j];
        *band_at(B, 3 * j, 3 * j - 1) = h * h * h;
        *band_at(B, 3 * j + 1, 3 * j - 1) = 3.0 * h * h;
        *band_at(B, 3 * j - 2, 3 * j) = 2.0;
        *band_at(B, 3 * j, 3 * j) = h * h;
        *band_at(B, 3 * j + 1, 3 * j) = 2.0 * h;
        *band_at(B, 3 * j - 1, 3 * j + 1) = 1.0;
        *band_at(B, 3 * j, 3 * j + 1) = h;
        *band_at(B, 3 * j + 1, 3 * j + 1) = 1.0;
        r[3 * j] = V[j + 1] - V[j];
        r[3 * j + 1] = dVN0;
    }

    /* banded LU with partial pivoting (dgbtf2), forward elimination of the rhs fused
       in the same column order as dgbtrs applies it */
    for (int j = 0; j < N; j++) {
        const int km = (SKL < N - 1 - j) ? SKL : N - 1 - j;
        int p = j;
        double amax = fabs(*band_at(B, j, j));
        for (int i = j + 1; i <= j + km; i++) {
            double v = fabs(*band_at(B, i, j));
            if (v > amax) {
                amax = v;
                p = i;
            }
        }
        const int cmax = (j + SKL + SKU < N - 1) ? j + SKL + SKU : N - 1;
        if (*band_at(B, p, j) != 0.0) {
            if (p != j) {
                for (int c = j; c <= cmax; c++) {
                    double t = *band_at(B, p, c);
                    *band_at(B, p, c) = *band_at(B, j, c);
                    *band_at(B, j, c) = t;
                }
                double t = r[p];
                r[p] = r[j];
                r[j] = t;
            }
            const double rp = 1.0 / *band_at(B, j, j);
            for (int i = j + 1; i <= j + km; i++) {
                const double l = *band_at(B, i, j) * rp;
                *band_at(B, i, j) = l;
                for (int c = j + 1; c <= cmax; c++) {
                    const double y = *band_at(B, j, c);
                    if (y != 0.0) *band_at(B, i, c) += l * (-y);
                }
                r[i] += l * (-r[j]);
            }
        }
    }
    /* back substitution with the upper band (KL+KU superdiagonals), column oriented */
    for (int j = N - 1; j >= 0; j--) {
        if (r[j] != 0.0) {
            r[j] = r[j] / *band_at(B, j, j);
            const double t = r[j];
            const int i0 = (j - SKL - SKU > 0) ? j - SKL - SKU : 0;
            for (int i = j - 1; i >= i0; i--) r[i] -= t * *band_at(B, i, j);
        }
    }

    coeffs[0] = r[0];
    coeffs[1] = r[1];
    coeffs[2] = dV1;
    coeffs[3] = V[0];
    for (int j = 1; j < NS; j++) {
        coeffs[4 * j] = r[3 * j - 1];
        coeffs[4 * j + 1] = r[3 * j];
        coeffs[4 * j + 2] = r[3 * j + 1];
        coeffs[4 * j + 3] = V[j];
    }
}

void orc_spline_coeffs_rel(const double *x0, const double *V, int N0, double *coeffs)
{
    const int N = 3 * (N0 - 1) - 1;
    double *B = (double *)orc_xmalloc(sizeof(double) * (size_t)N * SW);
    double *r = (double *)orc_xmalloc(sizeof(double) * (size_t)N);
    orc_spline_coeffs_rel_ws(x0, V, N0, coeffs, B, r);
    free(B);
    free(r);
}

/* ------------------------------------------------------------------------------------
 * Root finder, generique.cpp:604-913 (lims == NULL: tol 1e-16, 20 iterations,
 * eps = 10 DBL_EPSILON) and generique.cpp:917-1234 (lims given: tol 1e-15,
 * 12 iterations, eps = 2 DBL_EPSILON, returns false if the bracket leaves lims).
 * Bracketed inverse-quadratic steps with secant fallback and bisection when progress
 * stalls; a secant (Newton-Raphson with a finite-difference slope) search first when
 * the initial interval does not bracket a sign change.  The final root is the abscissa
 * of the smallest |f| among the three kept points.
 * The reference leaves the Newton start point unset when f(init[0]) already has the
 * smallest |f|; here it starts from init[0].
 * ---------------------------------------------------------------------------------- */
static void pick_sides(const double *x, const double *y, int *ind3, int *indmid, int *indfar,
                       int *indm, int *indp)
{
    int ind1, ind2;
    if (y[0] * y[1] > 0)
        ind1 = 0, ind2 = 1;
    else if (y[0] * y[2] > 0)
        ind1 = 0, ind2 = 2;
    else
        ind1 = 1, ind2 = 2;
    *ind3 = 3 - ind1 - ind2;
    if (x[*ind3] > x[ind1]) {
        if (x[ind1] > x[ind2])
            *indfar = ind2, *indmid = ind1;
        else
            *indfar = ind1, *indmid = ind2;
        *indm = *indmid, *indp = *ind3;
    } else {
        if (x[ind1] < x[ind2])
            *indfar = ind2, *indmid = ind1;
        else
            *indfar = ind1, *indmid = ind2;
        *indm = *ind3, *indp = *indmid;
    }
}

int orc_find_zero(orc_fn f, void *ctx, const double init[3], double *root, const double *lims,
                  int *nevals)
{
    const double tol = lims ? 1.0e-15 : 1.0e-16;
    const double eps = (lims ? 2.0 : 10.0) * DBL_EPSILON;
    const int itmax = lims ? 12 : 20;
    const double dytol = 0.01;
    int it = 0, ne = 0;
    double a = 0.0, b = 0.0, c = 0.0;
    const double x0 = init[0], xmin = init[1], xmax = init[2];
    double ymin = f(ctx, xmin), ymax = f(ctx, xmax);
    ne += 2;
    double x[3] = {x0, xmin, xmax}, y[3] = {0.0, ymin, ymax};
    int ok = 1;
#define FZ_RET(v, okv)           \
    do {                         \
        *root = (v);             \
        if (nevals) *nevals = ne; \
        return (okv);            \
    } while (0)

    if (fabs(ymin) < tol) FZ_RET(xmin, 1);
    if (fabs(ymax) < tol) FZ_RET(xmax, 1);

    if (ymin * ymax > 0) {
        y[0] = f(ctx, x[0]);
        ne++;
        if (fabs(y[0]) < tol) FZ_RET(x[0], 1);
        if (y[0] * y[2] < 0) {
            double xt = x[0];
            x[0] = (x[0] + x[2]) / 2;
            x[1] = xt;
            y[1] = y[0];
        } else {
            double yt = y[0], xt = x[0], xt1, yt1;
            if (fabs(y[1]) < fabs(yt)) yt = y[1], xt = x[1];
            if (fabs(y[2]) < fabs(yt)) yt = y[2], xt = x[2];
            xt1 = xt + 10 * eps;
            yt1 = f(ctx, xt1);
            ne++;
            xt1 = xt - yt * ((xt1 - xt) / (yt1 - yt));
            yt1 = f(ctx, xt1);
            ne++;
            it++;
            while (fabs(yt1) > tol && it < itmax && yt1 * yt > 0 && isfinite(xt1)) {
                xt = xt1;
                yt = yt1;
                xt1 = xt + 10 * eps;
                yt1 = f(ctx, xt1);
                ne++;
                xt1 = xt - yt * ((xt1 - xt) / (yt1 - yt));
                yt1 = f(ctx, xt1);
                ne++;
                it++;
            }
            if (fabs(yt1) < tol) FZ_RET(xt1, 1);
            if (yt1 * yt < 0) {
                x[0] = (xt + xt1) / 2;
                x[1] = xt;
                y[1] = yt;
                x[2] = xt1;
                y[2] = yt1;
            }
            if (it >= itmax) FZ_RET(xt1, 0);
            if (!isfinite(xt1)) FZ_RET(xt1, 0);
        }
    }
    if (lims && (x[1] < lims[0] || x[1] > lims[1] || x[2] < lims[0] || x[2] > lims[1]))
        FZ_RET(x[0], 0);
    if (fabs(xmax - xmin) < eps) FZ_RET(xmin, 1);

    y[0] = f(ctx, x[0]);
    ne++;
    if (fabs(y[0]) < tol) FZ_RET(x[0], 1);

    int ind3, indmid, indfar, indm, indp;
    pick_sides(x, y, &ind3, &indmid, &indfar, &indm, &indp);
    x[indfar] = -y[indmid] * (x[ind3] - x[indmid]) / (y[ind3] - y[indmid]) + x[indmid];
    y[indfar] = f(ctx, x[indfar]);
    ne++;
    double dy = y[indfar], dyprec;

    while (fabs(dy) > tol && it < itmax) {
        pick_sides(x, y, &ind3, &indmid, &indfar, &indm, &indp);
        const double denom = -x[0] * x[0] * x[2] + x[0] * x[0] * x[1] - x[0] * x[1] * x[1] +
                             x[0] * x[2] * x[2] - x[2] * x[2] * x[1] + x[2] * x[1] * x[1];
        if (fabs(denom) > 0.0) {
            a = (x[0] * y[2] - x[0] * y[1] + x[2] * y[1] - y[2] * x[1] - y[0] * x[2] + y[0] * x[1]) / denom;
            b = -(-y[2] * x[1] * x[1] + x[1] * x[1] * y[0] - x[2] * x[2] * y[0] - y[1] * x[0] * x[0] +
                  y[2] * x[0] * x[0] + y[1] * x[2] * x[2]) / denom;
            c = (y[2] * x[0] * x[0] * x[1] - y[1] * x[0] * x[0] * x[2] - x[1] * x[1] * x[0] * y[2] +
                 x[2] * x[2] * x[0] * y[1] + x[1] * x[1] * y[0] * x[2] - x[2] * x[2] * y[0] * x[1]) / denom;
        }
        if ((b * b - 4.0 * a * c) > 0.0 && fabs(denom) > 0.0) {
            const double xp = (-b + sqrt(b * b - 4.0 * a * c)) / (2 * a);
            const double xm = (-b - sqrt(b * b - 4.0 * a * c)) / (2 * a);
            if (xp > x[indm] && xp < x[indp])
                x[indfar] = xp;
            else if (xm > x[indm] && xm < x[indp])
                x[indfar] = xm;
            else if (fabs(y[ind3] - y[indmid]) > tol)
                x[indfar] = -y[indmid] * (x[ind3] - x[indmid]) / (y[ind3] - y[indmid]) + x[indmid];
            else
                FZ_RET(x[indmid], 1);
        } else {
            if (fabs(y[ind3] - y[indmid]) > tol)
                x[indfar] = -y[indmid] * (x[ind3] - x[indmid]) / (y[ind3] - y[indmid]) + x[indmid];
            else
                FZ_RET(x[indmid], 1);
        }
        y[indfar] = f(ctx, x[indfar]);
        ne++;
        dyprec = fabs(dy);
        dy = y[indfar];
        if (fabs(dyprec - fabs(dy)) < dytol * dyprec || dyprec < fabs(dy)) {
            x[indfar] = (x[ind3] + x[indmid]) / 2.0;
            y[indfar] = f(ctx, x[indfar]);
            ne++;
            dyprec = fabs(dy);
            dy = y[indfar];
        }
        it++;
    }
    (void)dyprec;
    int best = 0;
    if (fabs(y[1]) < fabs(y[best])) best = 1;
    if (fabs(y[2]) < fabs(y[best])) best = 2;
    (void)ok;
    FZ_RET(x[best], 1);
#undef FZ_RET
}

/* ------------------------------------------------------------------------------------
 * Roots of A z^2 + B z + C with Newton polishing, generique.cpp:1666-1726.
 * ---------------------------------------------------------------------------------- */
static double cabs_ref(double complex z) { return sqrt(creal(z) * creal(z) + cimag(z) * cimag(z)); }

void orc_roots_2pol(const double complex coef[3], double complex roots[2])
{
    const double complex A = coef[0], B = coef[1], C = coef[2];
    if (!(cabs_ref(A) > 1.0e-15)) return;
    const double complex R = csqrt(B * B - 4.0 * A * C);
    roots[0] = (-B - R) / (2.0 * A);
    roots[1] = (-B + R) / (2.0 * A);
    for (int j = 0; j < 2; j++) {
        double complex x = roots[j];
        double complex yv = A * x * x + B * x + C;
        double complex p = 2.0 * A * x + B;
        const double complex y0 = yv;
        int it = 0;
        while (cabs_ref(p) != 0.0 && it < 10 && cabs_ref(yv) > 1.0e-16) {
            const double complex x2 = x, y2 = yv;
            const double complex cc = yv - p * x;
            x = -cc / p;
            yv = A * x * x + B * x + C;
            p = 2.0 * A * x + B;
            if (cabs_ref(y2) < cabs_ref(yv)) {
                x = 0.5 * (x + x2);
                yv = A * x * x + B * x + C;
                p = 2.0 * A * x + B;
            }
            it++;
        }
        if (cabs_ref(yv) < cabs_ref(y0)) roots[j] = x;
    }
}

/* ------------------------------------------------------------------------------------
 * Thiele continued-fraction Pade, generique.h:112 + generique.cpp:1630-1663.
 * coef[j] = g_j(x_j) with g_0 = func and g_j(x) = (g_{j-1}(x_{j-1}) - g_{j-1}(x)) /
 * ((x - x_{j-1}) g_{j-1}(x)); evaluated with the forward A/B recurrence of pade().
 * ---------------------------------------------------------------------------------- */
void orc_pade_coef(const double complex *func, const double complex *x, int N, double complex *coef)
{
    double complex *g = (double complex *)orc_xmalloc(sizeof(double complex) * (size_t)N);
    for (int j = 0; j < N; j++) {
        /* g holds g_{j}(x_i) for i >= j after the update below */
        if (j == 0) {
            for (int i = 0; i < N; i++) g[i] = func[i];
        } else {
            for (int i = j; i < N; i++) g[i] = (coef[j - 1] - g[i]) / ((x[i] - x[j - 1]) * g[i]);
        }
        coef[j] = g[j];
    }
    free(g);
}

double complex orc_pade_eval(double complex x, int n, const double complex *x0, const double complex *coef)
{
    double complex A = 1.0, B = 1.0 + (x - x0[0]) * coef[1];
    double complex q = A / B;
    for (int j = 3; j <= n; j++) {
        A = 1.0 + (x - x0[j - 2]) * coef[j - 1] / A;
        B = 1.0 + (x - x0[j - 2]) * coef[j - 1] / B;
        q = q * A / B;
    }
    return q * coef[0];
}

void orc_pade_eval_out(const double z[2], int N, const double complex *x0, const double complex *coef,
                       double out[2])
{
    const double complex v = orc_pade_eval(z[0] + I * z[1], N, x0, coef);
    out[0] = creal(v);
    out[1] = cimag(v);
}
