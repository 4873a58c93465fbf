/*
 * tpsc_oracle.c -- FP64 CPU restatement of the TPSC hot path (test infrastructure only;
 * see tpsc_oracle.h).  Stage by stage it follows dbergeron1/TPSC_conductivity:
 *   hamiltonien.cpp:14-34,49-68,119-155   dispersion, real-space hoppings
 *   green0.cpp:182-262                    finite-lattice mu0
 *   chi0.cpp:279-718                      chi0(q,tau) slabs, chi0(q,iqn) pencils
 *   chi.cpp:37-206,210-417,738-876        V(r,tau), Usp/Uch sum rules
 *   green.cpp:1673-2031,2929-3282         Sigma(k,iwn), traces, interacting mu
 *   cond_opt.cpp:2537-4290                Lambda_jj bubble + Maki-Thompson + kx0
 * Reductions use a fixed block decomposition so results do not depend on the thread count.
 */
#include "tpsc_oracle.h"

#include <omp.h>

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.14159265358979323846
#define NBLK 64

void orc_spline_coeffs_rel_ws(const double *x0, const double *V, int N0, double *coeffs,
                              double *B, double *r);

static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "tpsc_oracle: out of memory (%zu bytes)\n", n);
        abort();
    }
    return p;
}
static double *dalloc(size_t n) { return (double *)xmalloc(n * sizeof(double)); }
static double *dcalloc(size_t n)
{
    double *p = dalloc(n);
    memset(p, 0, n * sizeof(double));
    return p;
}
static void dfree(double **p)
{
    free(*p);
    *p = NULL;
}

static inline long widx(long l, long m) { return m + (l * (l + 1)) / 2; }
/* fold a lattice point onto the 1/8 wedge (green.cpp:1467-1476, cond_opt.cpp:2908-2917) */
static inline long fold_idx(long x, long y, long N)
{
    if (x > N / 2) x = N - x;
    if (y > N / 2) y = N - y;
    if (y > x) {
        long t = x;
        x = y;
        y = t;
    }
    return widx(x, y);
}
/* spline workspace sized for N0 points */
typedef struct {
    double *B, *r;
} sws;
static sws sws_new(int N0)
{
    sws w;
    long N = 3L * (N0 - 1) - 1;
    w.B = dalloc((size_t)N * 9);
    w.r = dalloc((size_t)N);
    return w;
}
static void sws_free(sws *w)
{
    free(w->B);
    free(w->r);
}

/* ---------------------------------------------------------------- model (a1) */
static inline double hop_t(const orc *o) { return o->p.t != 0.0 ? o->p.t : 1.0; }
double orc_dispk(const orc *o, double kx, double ky)
{
    const double tp = o->p.tp, tpp = o->p.tpp, t3 = o->p.t3, t = hop_t(o);
    return -2.0 * t * (cos(kx) + cos(ky)) - 4.0 * tp * cos(kx) * cos(ky) - 2.0 * tpp * (cos(2 * kx) + cos(2 * ky)) -
           4 * t3 * (cos(2 * kx) * cos(ky) + cos(2 * ky) * cos(kx));
}
double orc_grad_ek_x(const orc *o, double kx, double ky)
{
    const double tp = o->p.tp, tpp = o->p.tpp, t3 = o->p.t3, t = hop_t(o);
    return 2.0 * t * sin(kx) + 4.0 * tp * sin(kx) * cos(ky) + 4.0 * tpp * sin(2 * kx) +
           4 * t3 * (2.0 * sin(2 * kx) * cos(ky) + cos(2 * ky) * sin(kx));
}
double orc_hess_ek_xx(const orc *o, double kx, double ky)
{
    const double tp = o->p.tp, tpp = o->p.tpp, t3 = o->p.t3, t = hop_t(o);
    return 2.0 * t * cos(kx) + 4.0 * tp * cos(kx) * cos(ky) + 8.0 * tpp * cos(2 * kx) +
           4 * t3 * (4.0 * cos(2 * kx) * cos(ky) + cos(2 * ky) * cos(kx));
}

/* hamiltonien.cpp:49-68 (calc_ek) and :167-184 (set_system_size) */
static void calc_ek(orc *o)
{
    const int nbk = o->nbk;
    free(o->ek);
    free(o->er);
    o->ek = dalloc((size_t)nbk * nbk);
    for (int l = 0; l < nbk; l++) {
        const double kx = l * PI / (nbk - 1);
        for (int m = 0; m < nbk; m++) {
            const double ky = m * PI / (nbk - 1);
            o->ek[m + l * nbk] = orc_dispk(o, kx, ky);
        }
    }
    const long nk8 = (long)nbk * (nbk + 1) / 2;
    o->er = dcalloc((size_t)(nk8 > 5 ? nk8 : 5));
    o->er[1] = -hop_t(o);
    o->er[2] = -o->p.tp;
    o->er[3] = -o->p.tpp;
    o->er[4] = -o->p.t3;
}

/* green0.cpp:182-203 */
double orc_ddensity_mu_PBC(orc *o, double mup)
{
    const int nbk = o->nbk;
    const double Nk2 = 4.0 * (nbk - 1) * (nbk - 1);
    double dens = 0;
    for (int l = 0; l < nbk; l++)
        for (int m = 0; m <= l; m++) {
            double w = 8;
            if (l == 0 || l == nbk - 1) w = w / 2;
            if (m == 0 || m == nbk - 1) w = w / 2;
            if (m == l) w = w / 2;
            dens += 2 * w / ((exp((o->ek[m + l * nbk] - mup) / o->p.T) + 1.0) * Nk2);
        }
    return dens - o->p.n;
}
static double fz_density(void *ctx, double x) { return orc_ddensity_mu_PBC((orc *)ctx, x); }

/* green0.cpp:240-262 */
int orc_find_mu0(orc *o)
{
    if (o->mu0_fixed) return 1;
    if (o->p.n == 1.0 && o->p.tp == 0 && o->p.tpp == 0) {
        o->mu0 = 0.0;
        return 1;
    }
    const double init[3] = {0.0, -8.0, 8.0};
    double root = 0.0;
    if (orc_find_zero(fz_density, o, init, &root, NULL, NULL)) {
        o->mu0 = root;
        return 1;
    }
    return 0;
}

/* green0.h:44 */
void orc_set_grid(orc *o, int nk0, int Nw0)
{
    o->nbk = nk0 / 2 + 1;
    calc_ek(o);
    o->nbw = Nw0;
    o->nwmax = Nw0 / 2;
    orc_find_mu0(o);
}

orc *orc_create(const orc_params *p)
{
    orc *o = (orc *)xmalloc(sizeof(orc));
    memset(o, 0, sizeof(*o));
    o->p = *p;
    o->nsr = (p->n > 1.0) ? 2.0 - p->n : p->n; /* chi.cpp:18-19 */
    o->N_theta = 1;
    orc_set_grid(o, p->nk0, p->Nw0);
    return o;
}

void orc_destroy(orc *o)
{
    if (!o) return;
    free(o->ek);
    free(o->er);
    free(o->chiqw);
    free(o->chiqt);
    free(o->chirt);
    free(o->dtau_chirt0);
    free(o->dtau_Grt0p);
    free(o->dtau_chiqt0);
    free(o->chi0_inf);
    free(o->chi0_inf2);
    free(o->V);
    free(o->Sigma_tau);
    free(o->dtau_Sigma);
    free(o->Self);
    free(o->Sigma_inf2);
    free(o->Sigma_inf3);
    free(o->Sigma_inf4);
    free(o->chijj1);
    free(o->chijj2);
    free(o->chijj1_k);
    free(o->chijj2_k);
    free(o->al_freq);
    free(o->chijj3);
    free(o->chijj_all);
    free(o->chijj3_k);
    free(o->al_hw);
    free(o->al_hg2);
    free(o);
}

/* fill a square nk x nk slab symmetrically from a packed-wedge array */
static void wedge_to_square(const double *w, double *sq, int nk)
{
    for (int l = 0; l < nk; l++)
        for (int m = 0; m <= l; m++) {
            const double v = w[widx(l, m)];
            sq[m + l * nk] = v;
            sq[l + m * nk] = v;
        }
}
static void square_to_wedge(const double *sq, double *w, int nk)
{
    for (int l = 0; l < nk; l++)
        for (int m = 0; m <= l; m++) w[widx(l, m)] = sq[m + l * nk];
}

/* ---------------------------------------------------------------- chi0 (a3, a4) */
int orc_chi0(orc *o, int nq0, int nw0)
{
    const int nk = nq0 / 2 + 1;
    const double nk02 = (double)nq0 * nq0;
    const int nk2 = nk * nk;
    const long nk8 = (long)nk * (nk + 1) / 2;
    const int nw = nw0 / 2 + 1;
    const double T = o->p.T;

    orc_set_grid(o, nq0, nw0); /* chi0.cpp:297 */
    const double mu0 = o->mu0;

    free(o->chiqw);
    free(o->chiqt);
    free(o->chirt);
    free(o->dtau_chirt0);
    free(o->dtau_Grt0p);
    free(o->dtau_chiqt0);
    free(o->chi0_inf);
    free(o->chi0_inf2);
    o->chiqw = dalloc((size_t)nw * nk8);
    o->chirt = dalloc((size_t)nw * nk8);
    o->dtau_chirt0 = dalloc(nk8);
    o->dtau_Grt0p = dalloc(nk8);
    o->dtau_chiqt0 = dalloc(nk8);
    o->chi0_inf = dalloc(nk8);
    o->chi0_inf2 = dalloc(nk8);

    double *ekmu = dalloc(nk8), *Gkt = dalloc(nk8), *Grt0 = dalloc(nk8), *dGm = dalloc(nk8);
    double *sq = dalloc(nk2), *wv = dalloc(nk8);
    for (int l = 0; l < nk; l++)
        for (int m = 0; m <= l; m++) {
            const long i = widx(l, m);
            ekmu[i] = o->ek[m + l * nk] - mu0;
            Gkt[i] = -1.0 / (exp(-ekmu[i] / T) + 1.0);
        }
    /* d/dtau G(r,tau) at 0+ and 0- (chi0.cpp:323-367) */
    dGm[0] = -mu0;
    for (long i = 1; i < nk8; i++) dGm[i] = o->er[i];
    for (long i = 0; i < nk8; i++) wv[i] = -ekmu[i] * Gkt[i] / nk02;
    wedge_to_square(wv, sq, nk);
    orc_r2r_2d(sq, nk, nk, 0, 0);
    square_to_wedge(sq, wv, nk);
    for (long i = 0; i < nk8; i++) {
        dGm[i] -= wv[i];
        o->dtau_Grt0p[i] = wv[i];
    }
    /* tau = 0 slice (chi0.cpp:375-418) */
    for (long i = 0; i < nk8; i++) wv[i] = Gkt[i] / nk02;
    wedge_to_square(wv, sq, nk);
    orc_r2r_2d(sq, nk, nk, 0, 0);
    square_to_wedge(sq, Grt0, nk);
    for (long i = 0; i < nk8; i++) wv[i] = -2.0 * Grt0[i] * Grt0[i];
    wv[0] = -2.0 * Grt0[0] * (Grt0[0] + 1.0);
    memcpy(o->chirt, wv, sizeof(double) * nk8);
    wedge_to_square(wv, sq, nk);
    orc_r2r_2d(sq, nk, nk, 0, 0);
    square_to_wedge(sq, o->chiqw, nk);
    /* d/dtau chi0 at tau=0 (chi0.cpp:420-447) */
    for (long i = 0; i < nk8; i++) wv[i] = -2.0 * (o->dtau_Grt0p[i] * Grt0[i] + dGm[i] * Grt0[i]);
    wv[0] = -2.0 * (o->dtau_Grt0p[0] * (1.0 + Grt0[0]) + dGm[0] * Grt0[0]);
    memcpy(o->dtau_chirt0, wv, sizeof(double) * nk8);
    wedge_to_square(wv, sq, nk);
    orc_r2r_2d(sq, nk, nk, 0, 0);
    square_to_wedge(sq, o->dtau_chiqt0, nk);

    /* 0 < tau < beta/2 (chi0.cpp:454-523) */
#pragma omp parallel
    {
        double *s = dalloc(nk2), *w1 = dalloc(nk8), *w2 = dalloc(nk8);
#pragma omp for schedule(static)
        for (int j = 1; j < nw - 1; j++) {
            const double t = j / (nw0 * T);
            for (long i = 0; i < nk8; i++) {
                const double x = ekmu[i];
                if (x >= 0.0)
                    w1[i] = exp(-t * x) * Gkt[i] / nk02;
                else
                    w1[i] = -exp((1.0 / T - t) * x) / (exp(x / T) + 1.0) / nk02;
            }
            wedge_to_square(w1, s, nk);
            orc_r2r_2d(s, nk, nk, 0, 0);
            square_to_wedge(s, w1, nk); /* G(r,t) */
            for (long i = 0; i < nk8; i++) {
                const double x = ekmu[i];
                if (x >= 0.0)
                    w2[i] = -exp((t - 1.0 / T) * x) * Gkt[i] / nk02;
                else
                    w2[i] = exp(t * x) / ((exp(x / T) + 1.0) * nk02);
            }
            wedge_to_square(w2, s, nk);
            orc_r2r_2d(s, nk, nk, 0, 0);
            square_to_wedge(s, w2, nk); /* G(r,-t) */
            for (long i = 0; i < nk8; i++) w2[i] = -2.0 * w2[i] * w1[i];
            memcpy(o->chirt + (size_t)j * nk8, w2, sizeof(double) * nk8);
            wedge_to_square(w2, s, nk);
            orc_r2r_2d(s, nk, nk, 0, 0);
            square_to_wedge(s, o->chiqw + (size_t)j * nk8, nk);
        }
        free(s);
        free(w1);
        free(w2);
    }
    /* tau = beta/2 (chi0.cpp:527-559) */
    {
        const int j = nw - 1;
        const double t = 1.0 / (2.0 * T);
        for (long i = 0; i < nk8; i++) {
            const double x = ekmu[i];
            if (x >= 0.0)
                wv[i] = exp(-t * x) * Gkt[i] / nk02;
            else
                wv[i] = -exp((1.0 / T - t) * x) / ((exp(x / T) + 1.0) * nk02);
        }
        wedge_to_square(wv, sq, nk);
        orc_r2r_2d(sq, nk, nk, 0, 0);
        for (int i = 0; i < nk2; i++) sq[i] = 2.0 * sq[i] * sq[i];
        square_to_wedge(sq, o->chirt + (size_t)j * nk8, nk);
        orc_r2r_2d(sq, nk, nk, 0, 0);
        square_to_wedge(sq, o->chiqw + (size_t)j * nk8, nk);
    }
    free(ekmu);
    free(Gkt);
    free(Grt0);
    free(dGm);
    free(sq);
    free(wv);

    o->chiqt = dalloc((size_t)nw * nk8);
    memcpy(o->chiqt, o->chiqw, sizeof(double) * (size_t)nw * nk8);

    /* tau -> iqn by clamped spline + DFT (chi0.cpp:566-698) */
    const int NS0 = nw;
#pragma omp parallel
    {
        double *tau = dalloc(NS0), *val = dalloc(NS0), *cf = dalloc(4 * (size_t)(NS0 - 1));
        double complex *z = (double complex *)xmalloc(sizeof(double complex) * (size_t)nw0);
        sws ws = sws_new(NS0);
#pragma omp for schedule(dynamic, 16)
        for (long q = 0; q < nk8; q++) {
            for (int j = 0; j < NS0; j++) {
                tau[j] = j / (nw0 * T);
                val[j] = o->chiqw[q + (size_t)j * nk8];
            }
            cf[0] = o->dtau_chiqt0[q];
            cf[1] = 0;
            orc_spline_coeffs_rel_ws(tau, val, NS0, cf, ws.B, ws.r);
            for (int j = 0; j < nw0 / 2; j++) z[j] = 6.0 * cf[4 * j];
            for (int j = nw0 / 2; j < nw0; j++) z[j] = -z[nw0 - j - 1];
            orc_dft(z, nw0, +1);
            double integ = 0;
            for (int j = 0; j < nw0 / 2; j++) {
                const double a = cf[4 * j], b = cf[4 * j + 1], c = cf[4 * j + 2], d = cf[4 * j + 3];
                const double x = tau[j + 1] - tau[j];
                integ += a * x * x * x * x / 4.0 + b * x * x * x / 3.0 + c * x * x / 2.0 + d * x;
            }
            o->chiqw[q] = 2 * integ;
            if (q != 0) {
                o->chi0_inf[q] = -2.0 * o->dtau_chiqt0[q];
                for (int j = 1; j < nw; j++) {
                    const double wn = 2 * j * PI * T;
                    o->chiqw[q + (size_t)j * nk8] =
                        o->chi0_inf[q] / (wn * wn) +
                        ((1.0 - cos(2 * j * PI / nw0)) * creal(z[j]) + sin(2 * j * PI / nw0) * cimag(z[j])) /
                            (wn * wn * wn * wn);
                }
                o->chi0_inf2[q] = 2.0 * creal(z[nw - 1]);
            } else {
                o->chi0_inf[0] = 0;
                o->chi0_inf2[0] = 0;
                for (int j = 1; j < nw; j++) o->chiqw[(size_t)j * nk8] = 0;
            }
        }
        free(tau);
        free(val);
        free(cf);
        free(z);
        sws_free(&ws);
    }
    /* chi0max (chi0.cpp:703-717) */
    o->chi0max = 0.0;
    o->qmax[0] = o->qmax[1] = 0;
    for (int l = nk - 1; l >= 0; l--)
        for (int m = l; m >= 0; m--)
            if (o->chiqw[widx(l, m)] > o->chi0max) {
                o->chi0max = o->chiqw[widx(l, m)];
                o->qmax[0] = l;
                o->qmax[1] = m;
            }
    return 0;
}

/* chi0.cpp:1588-1651 */
static double chi0_integrand(orc *o, double kx, double ky, double qx, double qy, double wn)
{
    const double T = o->p.T, mu0 = o->mu0;
    const double e = orc_dispk(o, kx, ky), ekq = orc_dispk(o, kx + qx, ky + qy);
    double de = e - ekq;
    if (qx == 0.0 && qy == 0.0) de = 0.0;
    const double expk = exp((e - mu0) / T);
    if (wn != 0.0) {
        if (de != 0.0) return (1.0 / (expk + 1.0) - 1.0 / (exp((ekq - mu0) / T) + 1.0)) * de / (wn * wn + de * de);
        return 0;
    } else if (fabs(de) > 1e-12)
        return (1.0 / (expk + 1.0) - 1.0 / (exp((ekq - mu0) / T) + 1.0)) / de;
    if ((e - mu0) < 0) return -(1.0 / T) * expk / ((expk + 1.0) * (expk + 1.0));
    const double expkm = exp(-(e - mu0) / T);
    return -(1.0 / T) * expkm / ((expkm + 1.0) * (expkm + 1.0));
}
double orc_chiqw_discret(orc *o, int qx, int qy, int n)
{
    const int nq0 = 2 * (o->nbk - 1);
    const double Qx = qx * 2 * PI / nq0, Qy = qy * 2 * PI / nq0, wn = 2 * n * PI * o->p.T;
    double s = 0;
    for (int j = -nq0 / 2; j < nq0 / 2; j++) {
        const double kx = j * 2 * PI / nq0;
        for (int l = -nq0 / 2; l < nq0 / 2; l++) s -= 2 * chi0_integrand(o, kx, l * 2 * PI / nq0, Qx, Qy, wn) / ((double)nq0 * nq0);
    }
    return s;
}

/* chi0.cpp:1498-1584 (nwr = 0) */
double orc_trace_chi0(orc *o, double *tr_corr)
{
    const int nbq = o->nbk, nwmax = o->nwmax;
    const long nk8 = (long)nbq * (nbq + 1) / 2;
    const double Nq = 4.0 * (nbq - 1) * (nbq - 1), T = o->p.T;
    const int N1 = nwmax + 1;
    double tr = 0.0;
    for (int j = 0; j < N1; j++)
        for (int l = 0; l < nbq; l++)
            for (int m = 0; m <= l; m++) {
                int w = 16;
                if (j == 0 || j == nwmax) w /= 2;
                if (l == 0 || l == nbq - 1) w /= 2;
                if (m == 0 || m == nbq - 1) w /= 2;
                if (m == l) w /= 2;
                tr += w * T * o->chiqw[widx(l, m) + (size_t)j * nk8] / Nq;
            }
    const double tr1 = tr;
    for (int j = nwmax + 1; j < 2 * nwmax; j++)
        for (int l = 0; l < nbq; l++)
            for (int m = 0; m <= l; m++) {
                int w = 16;
                if (l == 0 || l == nbq - 1) w /= 2;
                if (m == 0 || m == nbq - 1) w /= 2;
                if (m == l) w /= 2;
                const double wn = 2 * PI * j * T;
                const long i = widx(l, m);
                const double c = o->chi0_inf[i] / (wn * wn) + o->chi0_inf2[i] / (wn * wn * wn * wn);
                tr += w * T * c / Nq;
            }
    if (tr_corr) *tr_corr = tr + N1 * (tr1 - tr) / (N1 - 2 * nwmax);
    return tr;
}

/* ---------------------------------------------------------------- sum rules (a5) */
/* chi.cpp:738-843: channel 0 = spin (sum_rule_sp), 1 = charge (sum_rule_ch) */
double orc_sum_rule(orc *o, int channel, double Ux)
{
    const int nbq = o->nbk, nwmax = o->nwmax;
    const long nk8 = (long)nbq * (nbq + 1) / 2;
    const double Nq = 4.0 * (nbq - 1) * (nbq - 1), T = o->p.T;
    const double s = channel == 0 ? -1.0 : 1.0; /* chi/(1 + s*U*chi/2) */
    double part[NBLK];
    const long nterm = (long)(2 * nwmax + 1) * nk8;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < NBLK; b++) {
        const long i0 = nterm * b / NBLK, i1 = nterm * (b + 1) / NBLK;
        double acc = 0;
        long l = 0, m = 0;
        /* decode (j, idx) from the flat index i = j*nk8 + idx */
        for (long i = i0; i < i1; i++) {
            const long j = i / nk8, idx = i % nk8;
            if (i == i0 || idx == 0) {
                /* recover (l,m) from idx */
                l = (long)((sqrt(8.0 * idx + 1.0) - 1.0) / 2.0);
                while (widx(l + 1, 0) <= idx) l++;
                while (widx(l, 0) > idx) l--;
                m = idx - widx(l, 0);
            } else {
                m++;
                if (m > l) {
                    l++;
                    m = 0;
                }
            }
            int w = 16;
            if (j == 0) w /= 2;
            if (l == 0 || l == nbq - 1) w /= 2;
            if (m == 0 || m == nbq - 1) w /= 2;
            if (m == l) w /= 2;
            double c;
            if (j <= nwmax)
                c = o->chiqw[idx + j * nk8];
            else {
                const double wn = 2 * j * PI * T;
                c = o->chi0_inf[idx] / (wn * wn) + o->chi0_inf2[idx] / (wn * wn * wn * wn);
            }
            acc += w * T / Nq * (c / (1 + s * Ux * c / 2.0) - c);
        }
        part[b] = acc;
    }
    double tr = 0;
    for (int b = 0; b < NBLK; b++) tr += part[b];
    const double n = o->p.n, nsr = o->nsr, U = o->p.U;
    if (channel == 0) return tr + n - n * n / 2.0 - nsr + Ux * nsr * nsr / (2.0 * U);
    return tr + n - n * n / 2.0 - (nsr + o->Usp * nsr * nsr / (2.0 * U) - nsr * nsr);
}
static double fz_sp(void *ctx, double x) { return orc_sum_rule((orc *)ctx, 0, x); }
static double fz_ch(void *ctx, double x) { return orc_sum_rule((orc *)ctx, 1, x); }

/* the Uch root on the chi0 already computed, given o->Usp (the charge sum rule reads Usp
   through the double occupancy; chi.cpp:866-869) */
static int orc_uch_root(orc *o, double *Uch)
{
    double root = 0.0;
    const double init[3] = {0.0, 0.0, 1000.0 * o->p.U};
    if (!orc_find_zero(fz_ch, o, init, &root, NULL, &o->nevals_ch)) return 0;
    *Uch = root;
    return 1;
}

/* chi.cpp:846-876 */
/* the Usp and Uch roots on the chi0 already computed (chi.cpp:853-869) */
static int orc_usp_uch_roots(orc *o, double Usc[2])
{
    const double U = o->p.U;
    double root = 0.0;
    double init[3] = {U / 2.0, 0.0, 2.0 / o->chi0max - 2 * DBL_EPSILON};
    int ok = 1;
    if (orc_find_zero(fz_sp, o, init, &root, NULL, &o->nevals_sp)) {
        Usc[0] = root;
        o->Usp = root;
    } else
        ok = 0;
    if (!orc_uch_root(o, &Usc[1])) ok = 0;
    return ok;
}

int orc_find_Usp_Uch_grid(orc *o, int nq0, int nw0, double Usc[2])
{
    orc_chi0(o, nq0, nw0);
    return orc_usp_uch_roots(o, Usc);
}

/* chi.cpp:210-417; with chi0_done the full-grid chi0 is taken as computed (the stage-timed solve) */
static int orc_find_Usp_Uch_impl(orc *o, int chi0_done)
{
    double Usc[2] = {0.0, 0.0};
    const int N = o->p.nk0, Nw = o->p.Nw0;
    o->Usp_coarse = o->Uch_coarse = 0.0;
    if (o->p.coarse_pass && !chi0_done) {
        orc_find_Usp_Uch_grid(o, N / 2, Nw / 2, Usc);
        o->Usp_coarse = Usc[0];
        o->Uch_coarse = Usc[1];
    }
    int ok = chi0_done ? orc_usp_uch_roots(o, Usc) : orc_find_Usp_Uch_grid(o, N, Nw, Usc);
    o->difRelsp = fabs((Usc[0] - o->Usp_coarse) / Usc[0]);
    o->difRelch = fabs((Usc[1] - o->Uch_coarse) / Usc[1]);
    o->Usp = Usc[0];
    o->Uch = Usc[1];
    const double U = o->p.U, nsr = o->nsr;
    if (o->p.n > 1.0)
        o->dblOcc = o->Usp / U * nsr * nsr / 4.0 + 1.0 - nsr;
    else
        o->dblOcc = o->Usp / U * nsr * nsr / 4.0;
    return ok;
}

int orc_find_Usp_Uch(orc *o) { return orc_find_Usp_Uch_impl(o, 0); }

/* ---------------------------------------------------------------- vertex (a6) */
/* chi.cpp:37-206 */
int orc_vertex(orc *o)
{
    const int nbq = o->nbk, nw = o->nwmax + 1, nq0 = 2 * (nbq - 1);
    const long nk8 = (long)nbq * (nbq + 1) / 2;
    const double U = o->p.U, Usp = o->Usp, Uch = o->Uch, T = o->p.T;
    const double normFact = T / ((double)nq0 * nq0);
    free(o->V);
    o->V = dalloc((size_t)nw * nk8);
#pragma omp parallel
    {
        double *s = dalloc((size_t)nbq * nbq), *w = dalloc(nk8);
#pragma omp for schedule(static)
        for (int j = 0; j < nw; j++) {
            for (long i = 0; i < nk8; i++) {
                const double c = o->chiqw[i + (size_t)j * nk8];
                const double csp = c / (1.0 - Usp * c / 2.0), cch = c / (1.0 + Uch * c / 2.0);
                w[i] = normFact * U * (3.0 * Usp * (csp - c) + Uch * (cch - c)) / 8.0;
            }
            wedge_to_square(w, s, nbq);
            orc_r2r_2d(s, nbq, nbq, 0, 0);
            square_to_wedge(s, o->V + (size_t)j * nk8, nbq);
        }
        free(s);
        free(w);
    }
#pragma omp parallel
    {
        double *v = dalloc(nw);
#pragma omp for schedule(static)
        for (long i = 0; i < nk8; i++) {
            for (int j = 0; j < nw; j++) v[j] = o->V[i + (size_t)j * nk8];
            orc_redft00_1d(v, nw);
            for (int j = 0; j < nw; j++) o->V[i + (size_t)j * nk8] = v[j];
        }
        free(v);
    }
    const double fact = U * (3.0 * Usp + Uch) / 8.0;
    for (long i = 0; i < (long)nw * nk8; i++) o->V[i] += fact * o->chirt[i];
    return 0;
}

/* ---------------------------------------------------------------- self-energy (a7, a8) */
/* green.cpp:1737-2031 */
int orc_self_energy(orc *o)
{
    const int nk = o->nbk, nk2 = nk * nk, nk0 = 2 * (nk - 1);
    const double nk02 = (double)nk0 * nk0;
    const long nk8 = (long)nk * (nk + 1) / 2;
    const int nw = o->nwmax + 1, nbw = 2 * (nw - 1);
    const double T = o->p.T, mu0 = o->mu0, U = o->p.U;
    const double fact = U * (3.0 * o->Usp + o->Uch) / 8.0;

    double *ekmu = dalloc(nk2);
    for (int i = 0; i < nk2; i++) ekmu[i] = o->ek[i] - mu0;
    free(o->dtau_Sigma);
    o->dtau_Sigma = dalloc(2 * (size_t)nk8);
    double *sq = dalloc(nk2), *g = dalloc(nk8), *w = dalloc(nk8);
    /* tau = 0 (green.cpp:1745-1777) */
    for (int i = 0; i < nk2; i++) sq[i] = -1.0 / (exp(-ekmu[i] / T) + 1.0) / nk02;
    orc_r2r_2d(sq, nk, nk, 0, 0);
    square_to_wedge(sq, g, nk);
    for (long i = 0; i < nk8; i++) w[i] = fact * o->dtau_chirt0[i] * g[i] + o->V[i] * o->dtau_Grt0p[i];
    wedge_to_square(w, sq, nk);
    orc_r2r_2d(sq, nk, nk, 0, 0);
    square_to_wedge(sq, o->dtau_Sigma, nk);
    /* tau = beta (green.cpp:1780-1817) */
    for (int i = 0; i < nk2; i++) sq[i] = -1.0 / (exp(ekmu[i] / T) + 1.0) / nk02;
    orc_r2r_2d(sq, nk, nk, 0, 0);
    square_to_wedge(sq, g, nk);
    for (long i = 0; i < nk8; i++) {
        const double xi = (i == 0) ? -mu0 : o->er[i];
        w[i] = -fact * o->dtau_chirt0[i] * g[i] + o->V[i] * (-o->dtau_Grt0p[i] + xi);
    }
    wedge_to_square(w, sq, nk);
    orc_r2r_2d(sq, nk, nk, 0, 0);
    square_to_wedge(sq, o->dtau_Sigma + nk8, nk);
    free(sq);
    free(g);
    free(w);

    /* Sigma(k,tau), tau_j = j beta/Nw, j = 0..Nw (green.cpp:1822-1892) */
    free(o->Sigma_tau);
    o->Sigma_tau = dalloc((size_t)(nbw + 1) * nk8);
#pragma omp parallel
    {
        double *s = dalloc(nk2), *wv = dalloc(nk8);
#pragma omp for schedule(static)
        for (int j = 0; j <= nbw; j++) {
            const double t = j / (nbw * T);
            for (int i = 0; i < nk2; i++) {
                const double x = ekmu[i];
                if (x > 0)
                    s[i] = -exp(-t * x) / (exp(-x / T) + 1.0) / nk02;
                else
                    s[i] = -exp((1.0 / T - t) * x) / (exp(x / T) + 1.0) / nk02;
            }
            orc_r2r_2d(s, nk, nk, 0, 0);
            const int jv = (j < nw) ? j : nbw - j; /* V(r, beta - tau) for tau > beta/2 */
            square_to_wedge(s, wv, nk);
            for (long i = 0; i < nk8; i++) wv[i] *= o->V[i + (size_t)jv * nk8];
            wedge_to_square(wv, s, nk);
            orc_r2r_2d(s, nk, nk, 0, 0);
            square_to_wedge(s, o->Sigma_tau + (size_t)j * nk8, nk);
        }
        free(s);
        free(wv);
    }
    free(ekmu);

    /* tau -> iwn by clamped spline + twiddled DFT (green.cpp:1899-2018) */
    free(o->Self);
    free(o->Sigma_inf2);
    free(o->Sigma_inf3);
    free(o->Sigma_inf4);
    o->Self = (double complex *)xmalloc(sizeof(double complex) * (size_t)nw * nk8);
    memset(o->Self, 0, sizeof(double complex) * (size_t)nw * nk8);
    o->Sigma_inf2 = dalloc(nk8);
    o->Sigma_inf3 = dalloc(nk8);
    o->Sigma_inf4 = dalloc(nk8);
    const int NS0 = nbw + 1;
    double sig_inf = 0;
#pragma omp parallel
    {
        double *tau = dalloc(NS0), *val = dalloc(NS0), *cf = dalloc(4 * (size_t)nbw);
        double complex *z = (double complex *)xmalloc(sizeof(double complex) * (size_t)nbw);
        sws ws = sws_new(NS0);
#pragma omp for schedule(dynamic, 16)
        for (long l = 0; l < nk8; l++) {
            for (int j = 0; j <= nbw; j++) {
                tau[j] = j / (nbw * T);
                val[j] = o->Sigma_tau[l + (size_t)j * nk8];
            }
            cf[0] = o->dtau_Sigma[l];
            cf[1] = o->dtau_Sigma[l + nk8];
            orc_spline_coeffs_rel_ws(tau, val, NS0, cf, ws.B, ws.r);
            const double d2S0 = 2 * cf[1];
            const double d2Sb = 6.0 * cf[4 * nbw - 4] / (nbw * T) + 2.0 * cf[4 * nbw - 3];
            for (int j = 0; j < nbw; j++) {
                const double ang = j * PI / nbw;
                z[j] = (cos(ang) + I * sin(ang)) * (6.0 * cf[4 * j]);
            }
            orc_dft(z, nbw, +1);
            const double c1 = val[0] + val[nbw];
            if (l == 0) sig_inf = -c1;
            o->Sigma_inf2[l] = o->dtau_Sigma[l] + o->dtau_Sigma[l + nk8];
            o->Sigma_inf3[l] = -(d2S0 + d2Sb);
            for (int j = 0; j < nbw / 2; j++) {
                const double ang = (2 * j + 1) * PI / nbw;
                o->Self[l + (size_t)j * nk8] = (1.0 - (cos(ang) + I * sin(ang))) * z[j];
            }
            o->Sigma_inf4[l] = creal(o->Self[l + (size_t)(nbw / 2 - 1) * nk8]);
        }
        free(tau);
        free(val);
        free(cf);
        free(z);
        sws_free(&ws);
    }
    o->Sigma_inf = sig_inf;
    return 0;
}

/* brute-force Sigma(k,iwn), green.cpp:1401-1494 (needs chi0(q,iqn) kept) */
double complex orc_self_kw_brute(orc *o, int kx, int ky, int n)
{
    const int nbk = o->nbk, nk0 = 2 * (nbk - 1), nbw = o->nbw;
    const double Nk = (double)nk0 * nk0, T = o->p.T, U = o->p.U;
    const long nk8 = (long)nbk * (nbk + 1) / 2;
    const double kn = (2 * n + 1) * PI * T;
    double complex s = 0.0;
    for (long j = 0; j < 2L * nbw - 1; j++) {
        double qn;
        long wn;
        if (j < nbw) {
            qn = 2.0 * j * PI * T;
            wn = j;
        } else {
            qn = 2.0 * (j - 2 * nbw + 1) * PI * T;
            wn = 2 * nbw - j - 1;
        }
        for (long l = 0; l < nk0; l++) {
            long p = l + kx;
            if (p > nk0) p -= nk0;
            if (p > nbk - 1) p = nk0 - p;
            for (long m = 0; m < nk0; m++) {
                long q = m + ky;
                if (q > nk0) q -= nk0;
                if (q > nbk - 1) q = nk0 - q;
                const double ekq = o->ek[q + p * nbk];
                const long i = fold_idx(l, m, nk0);
                double c;
                if (wn <= nbw / 2)
                    c = o->chiqw[i + wn * nk8];
                else
                    c = o->chi0_inf[i] / (qn * qn) + o->chi0_inf2[i] / (qn * qn * qn * qn);
                const double csp = c / (1.0 - o->Usp * c / 2.0), cch = c / (1.0 + o->Uch * c / 2.0);
                const double complex G = 1.0 / ((-ekq + o->mu0) + I * (kn + qn));
                s += (U / 8.0) * (T / Nk) * (3.0 * o->Usp * csp + o->Uch * cch) * G;
            }
        }
    }
    return s;
}

/* Sigma(k, iwn) from its moment representation, n >= 0 (green.cpp:3039) */
static inline double complex sigma_at(const orc *o, long i, long j, double wn)
{
    const double complex z = I * wn;
    const long nk8 = (long)o->nbk * (o->nbk + 1) / 2;
    return o->Sigma_inf / z + o->Sigma_inf2[i] / (z * z) + o->Sigma_inf3[i] / (z * z * z) +
           o->Self[i + j * nk8] / (z * z * z * z);
}

static inline int wedge_w8(int l, int m, int nk)
{
    int w = 8;
    if (l == 0 || l == nk - 1) w /= 2;
    if (m == 0 || m == nk - 1) w /= 2;
    if (m == l) w /= 2;
    return w;
}

/* ---------------------------------------------------------------- traces, mu (a9) */
/* green.cpp:3001-3104 */
int orc_trace_self_G(orc *o)
{
    const int nk = o->nbk, nbw = o->nbw;
    const double nk02 = 4.0 * (nk - 1) * (nk - 1), T = o->p.T, mu0 = o->mu0;
    const int wl = nbw / 2;
    double sum = 0, sum2 = 0;
    for (int l = 0; l < nk; l++)
        for (int m = 0; m <= l; m++) {
            const double e = o->ek[m + l * nk];
            const int w = wedge_w8(l, m, nk);
            const long i = widx(l, m);
            for (int j = 0; j < wl; j++) {
                const double wn = (2 * j + 1) * PI * T;
                const double complex s = sigma_at(o, i, j, wn);
                const double tmp = w * T * 2 / (nk02 * (wn * wn + (e - mu0) * (e - mu0)));
                sum += tmp * (creal(s) * (mu0 - e) + cimag(s) * wn);
            }
        }
    for (int l = 0; l < nk; l++)
        for (int m = 0; m <= l; m++) {
            const double e = o->ek[m + l * nk];
            const int w = wedge_w8(l, m, nk);
            for (int j = 0; j < wl; j++) {
                const double wn = (2 * j + 1) * PI * T;
                sum2 += w * o->Sigma_inf * T * 2 / (nk02 * (wn * wn + (e - mu0) * (e - mu0)));
            }
            if (fabs(e - mu0) > 1.0e-12)
                sum2 += w * o->Sigma_inf * (1.0 / (exp((e - mu0) / T) + 1.0) - 0.5) / (e - mu0) / nk02;
            else {
                const double ek = (e - mu0) < 0 ? exp((e - mu0) / T) : exp(-(e - mu0) / T);
                sum2 -= w * o->Sigma_inf * ek / ((ek + 1.0) * (ek + 1.0) * nk02 * T);
            }
        }
    o->trSG = sum;
    o->trSG_corr = sum + sum2;
    return 0;
}

/* green.cpp:3108-3182 */
double orc_traceDiffG(orc *o, double pmu)
{
    const int nk = o->nbk, nbw = o->nbw;
    const double T = o->p.T, mu0 = o->mu0;
    double part[NBLK];
#pragma omp parallel for schedule(static)
    for (int b = 0; b < NBLK; b++) {
        double acc = 0;
        for (int l = b; l < nk; l += NBLK)
            for (int m = 0; m <= l; m++) {
                const double e = o->ek[m + l * nk];
                const double w = 2.0 * wedge_w8(l, m, nk);
                const long i = widx(l, m);
                for (int j = nbw; j >= nbw / 2; j--) {
                    const double wn = (2 * j + 1) * PI * T;
                    const double rS = -o->Sigma_inf2[i] / (wn * wn) + o->Sigma_inf4[i] / (wn * wn * wn * wn);
                    const double iS = -o->Sigma_inf / wn + o->Sigma_inf3[i] / (wn * wn * wn);
                    const double tmp = w / (((wn - iS) * (wn - iS) + (e - pmu + rS) * (e - pmu + rS)) *
                                            (wn * wn + (e - mu0) * (e - mu0)));
                    acc += tmp * ((mu0 - pmu + rS) * (e - pmu + rS) * (e - mu0) + iS * (iS - 2 * wn) * (e - mu0) -
                                  wn * wn * (mu0 - pmu + rS));
                }
                for (int j = nbw / 2 - 1; j >= 0; j--) {
                    const double wn = (2 * j + 1) * PI * T;
                    const double complex s = sigma_at(o, i, j, wn);
                    const double rS = creal(s), iS = cimag(s);
                    const double tmp = w / (((wn - iS) * (wn - iS) + (e - pmu + rS) * (e - pmu + rS)) *
                                            (wn * wn + (e - mu0) * (e - mu0)));
                    acc += tmp * ((mu0 - pmu + rS) * (e - pmu + rS) * (e - mu0) + iS * (iS - 2 * wn) * (e - mu0) -
                                  wn * wn * (mu0 - pmu + rS));
                }
            }
        part[b] = acc;
    }
    double s = 0;
    for (int b = 0; b < NBLK; b++) s += part[b];
    return s;
}
static double fz_mu(void *ctx, double x) { return orc_traceDiffG((orc *)ctx, x); }

/* green.cpp:3185-3282 */
int orc_find_mu(orc *o)
{
    const double init[3] = {0.0, -10.0, 10.0}, lims[2] = {-20, 20};
    double root = 0.0;
    if (orc_find_zero(fz_mu, o, init, &root, lims, &o->nevals_mu)) {
        o->mu = root;
        o->mu_ok = 1;
    } else
        o->mu_ok = 0;
    orc_trace_self_G2(o);
    return o->mu_ok;
}

/* green.cpp:2929-2997 */
int orc_trace_self_G2(orc *o)
{
    const int nk = o->nbk, nbw = o->nbw;
    const double nk02 = 4.0 * (nk - 1) * (nk - 1), T = o->p.T, mu = o->mu, U = o->p.U;
    double sum = 0;
    for (int l = 0; l < nk; l++)
        for (int m = 0; m <= l; m++) {
            const double e = o->ek[m + l * nk];
            const int w = wedge_w8(l, m, nk);
            const long i = widx(l, m);
            for (int j = 0; j < nbw / 2; j++) {
                const double wn = (2 * j + 1) * PI * T;
                const double complex s = sigma_at(o, i, j, wn);
                const double rS = creal(s), iS = cimag(s);
                const double tmp = w * T * 2 / (nk02 * ((wn - iS) * (wn - iS) + (e - mu + rS) * (e - mu + rS)));
                sum += tmp * (iS * wn - iS * iS - (e - mu) * rS - rS * rS);
            }
        }
    o->trSG2 = sum;
    o->dblOcc2 = (sum + U * o->p.n * o->p.n / 4.0) / U;
    return 0;
}

/* ---------------------------------------------------------------- Lambda_jj (a10) */
/* per-k Dyson G on the Matsubara grid ordered n = j - nw, j = 0..nbw-1; fills
   Gm[j] = G - G_asym and optionally J[j] = G*J - G_asym*Jinf (cond_opt.cpp:3354-3407) */
typedef struct {
    double complex r1, r2, a1, a2, b1, b2, den1, den2;
} pole_pair;

static void asym_poles(const orc *o, double ekv, double CJ1, pole_pair *pp)
{
    const double T = o->p.T;
    double complex coef[3] = {1.0, -ekv + o->mu, -o->Sigma_inf};
    double complex roots[2] = {0, 0};
    orc_roots_2pol(coef, roots);
    pp->r1 = roots[0];
    pp->r2 = roots[1];
    pp->a1 = pp->r1 / (pp->r1 - pp->r2);
    pp->a2 = pp->r2 / (pp->r2 - pp->r1);
    pp->b1 = CJ1 / (pp->r1 - pp->r2);
    pp->b2 = -pp->b1;
    pp->den1 = (creal(pp->r1) > 0 ? cexp(-pp->r1 / T) : cexp(pp->r1 / T)) + 1.0;
    pp->den2 = (creal(pp->r2) > 0 ? cexp(-pp->r2 / T) : cexp(pp->r2 / T)) + 1.0;
}

/* cond_opt.cpp:2537-4290 with chi3 = false */
int orc_chijj(orc *o, int N_theta)
{
    const int nbk = o->nbk, nbw = o->nbw, nw = nbw / 2, nk0 = 2 * (nbk - 1);
    const long nq = (long)nbk * (nbk - 2), nk8 = (long)nbk * (nbk + 1) / 2;
    const double T = o->p.T, mu0 = o->mu0, mu = o->mu, U = o->p.U;
    const double normFact = 1.0 / ((double)nk0 * nk0);
    const double *V = o->V;
    if (N_theta < 1) N_theta = 1;
    o->N_theta = N_theta;

    double *dek = dalloc((size_t)nq), *d2ek = dalloc((size_t)nbk * nbk), *ekmu = dalloc((size_t)nbk * nbk);
    for (int i = 0; i < nbk * nbk; i++) ekmu[i] = o->ek[i] - mu0;
    for (int l = 0; l < nbk; l++)
        for (int m = 0; m < nbk; m++) d2ek[m + l * nbk] = orc_hess_ek_xx(o, l * PI / (nbk - 1), m * PI / (nbk - 1));
    for (int l = 0; l < nbk - 2; l++)
        for (int m = 0; m < nbk; m++) dek[m + l * nbk] = orc_grad_ek_x(o, (l + 1) * PI / (nbk - 1), m * PI / (nbk - 1));

    const size_t hsz = (size_t)nq * (nbw + 1);
    double *h = dalloc(hsz), *f0 = dalloc(hsz);
    /* h(r,tau) (cond_opt.cpp:2697-2735) and f0(r,tau) (:2795-2827) */
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= nbw; j++) {
        const double t = (double)j / (T * nbw);
        double *fh = h + (size_t)j * nq, *ff = f0 + (size_t)j * nq;
        for (int l = 0; l < nbk - 2; l++)
            for (int m = 0; m < nbk; m++) {
                const double x = ekmu[m + (l + 1) * nbk], v = dek[m + l * nbk];
                const double e = o->ek[m + (l + 1) * nbk];
                if (x > 0)
                    fh[m + l * nbk] = normFact * v * exp(-x * t) / (exp(-x / T) + 1.0);
                else
                    fh[m + l * nbk] = normFact * v * exp(x * (1.0 / T - t)) / (exp(x / T) + 1.0);
                if (e > mu0) {
                    const double ex = exp((mu0 - e) / T);
                    ff[m + l * nbk] = normFact * v * exp((mu0 - e) * t) * (t / (ex + 1.0) - ex / (T * (ex + 1.0) * (ex + 1.0)));
                } else {
                    const double ex = exp((e - mu0) / T);
                    ff[m + l * nbk] = normFact * v * exp((e - mu0) * (1.0 / T - t)) * (t / (ex + 1.0) - 1.0 / (T * (ex + 1.0) * (ex + 1.0)));
                }
            }
        orc_r2r_2d(fh, nbk - 2, nbk, 1, 0);
        orc_r2r_2d(ff, nbk - 2, nbk, 1, 0);
    }
    /* boundary derivatives dh, df0 at tau = 0, beta (cond_opt.cpp:2745-2887) */
    double *dh = dalloc(2 * (size_t)nq), *df0 = dalloc(2 * (size_t)nq);
    for (int l = 0; l < nbk - 2; l++)
        for (int m = 0; m < nbk; m++) {
            const long q = m + l * nbk;
            const double e = o->ek[m + (l + 1) * nbk], v = dek[q];
            if (e > mu0) {
                const double ex = exp((mu0 - e) / T);
                dh[q] = -normFact * v * (e - mu0) / (ex + 1.0);
                dh[q + nq] = -normFact * v * (e - mu0) * ex / (ex + 1.0);
                df0[q] = normFact * v * ((e - mu0) * ex / (T * (ex + 1.0) * (ex + 1.0)) + 1.0 / (ex + 1.0));
                df0[q + nq] = normFact * v * ex * ((mu0 - e) * (1.0 / (T * (ex + 1.0)) - ex / (T * (ex + 1.0) * (ex + 1.0))) + 1.0 / (ex + 1.0));
            } else {
                const double ex = exp((e - mu0) / T);
                dh[q] = -normFact * v * (e - mu0) * ex / (ex + 1.0);
                dh[q + nq] = -normFact * v * (e - mu0) / (ex + 1.0);
                df0[q] = normFact * v * ex * ((e - mu0) / (T * (ex + 1.0) * (ex + 1.0)) + 1.0 / (ex + 1.0));
                df0[q + nq] = normFact * v * ((mu0 - e) * (1.0 / (T * (ex + 1.0)) - 1.0 / (T * (ex + 1.0) * (ex + 1.0))) + 1.0 / (ex + 1.0));
            }
        }
    orc_r2r_2d(dh, nbk - 2, nbk, 1, 0);
    orc_r2r_2d(dh + nq, nbk - 2, nbk, 1, 0);
    orc_r2r_2d(df0, nbk - 2, nbk, 1, 0);
    orc_r2r_2d(df0 + nq, nbk - 2, nbk, 1, 0);

    /* dhV, df0V (cond_opt.cpp:2892-3014); V_array is 2x vertex_rt_array */
    const double fact4 = U * (3.0 * o->Usp + o->Uch) / 4;
    double *dhV = dalloc(2 * (size_t)nq), *df0V = dalloc(2 * (size_t)nq);
    for (int l = 0; l < nbk - 2; l++)
        for (int m = 0; m < nbk; m++) {
            const long q = m + l * nbk, r = fold_idx(l + 1, m, nk0);
            dhV[q] = 2.0 * dh[q] * V[r] + h[q] * fact4 * o->dtau_chirt0[r];
            dhV[q + nq] = 2.0 * dh[q + nq] * V[r] - h[q + (size_t)nbw * nq] * fact4 * o->dtau_chirt0[r];
            df0V[q] = 2.0 * df0[q] * V[r] + f0[q] * fact4 * o->dtau_chirt0[r];
            df0V[q + nq] = 2.0 * df0[q + nq] * V[r] - f0[q + (size_t)nbw * nq] * fact4 * o->dtau_chirt0[r];
        }
    orc_r2r_2d(dhV, nbk - 2, nbk, 1, 0);
    orc_r2r_2d(dhV + nq, nbk - 2, nbk, 1, 0);
    orc_r2r_2d(df0V, nbk - 2, nbk, 1, 0);
    orc_r2r_2d(df0V + nq, nbk - 2, nbk, 1, 0);
    free(dh);
    free(df0);

    /* hV(k,tau), f0V(k,tau) overwrite h, f0 (cond_opt.cpp:3029-3114) */
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= nbw; j++) {
        const int jv = (j > nbw / 2) ? nbw - j : j;
        double *fh = h + (size_t)j * nq, *ff = f0 + (size_t)j * nq;
        for (int l = 0; l < nbk - 2; l++)
            for (int m = 0; m < nbk; m++) {
                const long q = m + l * nbk, r = fold_idx(l + 1, m, nk0) + (long)jv * nk8;
                fh[q] = 2.0 * fh[q] * V[r];
                ff[q] = 2.0 * ff[q] * V[r];
            }
        orc_r2r_2d(fh, nbk - 2, nbk, 1, 0);
        orc_r2r_2d(ff, nbk - 2, nbk, 1, 0);
    }

    /* per-k pencils (cond_opt.cpp:3179-3871), deterministic block partial sums */
    const int NS0 = nbw / 2 + 1;
    const int nwb = nw + 1;
    const double Dtheta = PI / (2.0 * N_theta);
    double *p1 = dcalloc((size_t)NBLK * nwb), *p2 = dcalloc((size_t)NBLK * nwb);
    double *p1k = dcalloc((size_t)NBLK * nwb * N_theta), *p2k = dcalloc((size_t)NBLK * nwb * N_theta);
    double pk0[NBLK], pinf[NBLK][3];
#pragma omp parallel
    {
        double complex *Jt = (double complex *)xmalloc(sizeof(double complex) * (size_t)nbw);
        double complex *J0 = (double complex *)xmalloc(sizeof(double complex) * (size_t)nbw);
        double complex *G = (double complex *)xmalloc(sizeof(double complex) * (size_t)nbw);
        double complex *GG = (double complex *)xmalloc(sizeof(double complex) * (size_t)nbw);
        double *tau2 = dalloc((size_t)nbw + 1), *v2 = dalloc((size_t)nbw + 1), *cf2 = dalloc(4 * (size_t)nbw);
        double *tau = dalloc(NS0), *v1 = dalloc(NS0), *cf = dalloc(4 * (size_t)(NS0 - 1));
        sws ws2 = sws_new(nbw + 1), ws1 = sws_new(NS0);
#pragma omp for schedule(static)
        for (int b = 0; b < NBLK; b++) {
            double *c1 = p1 + (size_t)b * nwb, *c2 = p2 + (size_t)b * nwb;
            double *c1k = p1k + (size_t)b * nwb * N_theta, *c2k = p2k + (size_t)b * nwb * N_theta;
            double k0acc = 0, inf1 = 0, inf2 = 0, inf4 = 0;
            const long qa = nq * b / NBLK, qb = nq * (b + 1) / NBLK;
            for (long qq = qa; qq < qb; qq++) {
                const int l = (int)(qq / nbk), m = (int)(qq % nbk);
                const long q = qq, ki = fold_idx(l + 1, m, nk0);
                const double C2 = o->Sigma_inf2[ki], C3 = o->Sigma_inf3[ki];
                const double ekv = o->ek[m + (l + 1) * nbk];
                const double v = dek[q];
                double CJ[2][3];
                /* splines + twiddled DFTs of hV and f0V -> J(ikn), J0(ikn) (:3276-3347) */
                for (int which = 0; which < 2; which++) {
                    const double *src = which == 0 ? h : f0;
                    const double *dsrc = which == 0 ? dhV : df0V;
                    double complex *dst = which == 0 ? Jt : J0;
                    for (int j = 0; j <= nbw; j++) {
                        tau2[j] = j / (nbw * T);
                        v2[j] = src[q + (size_t)j * nq];
                    }
                    cf2[0] = dsrc[q];
                    cf2[1] = dsrc[q + nq];
                    orc_spline_coeffs_rel_ws(tau2, v2, nbw + 1, cf2, ws2.B, ws2.r);
                    for (int j = 0; j < nbw; j++) {
                        const double ang = j * PI / nbw;
                        G[j] = (cos(ang) + I * sin(ang)) * (6.0 * cf2[4 * j]);
                    }
                    orc_dft(G, nbw, +1);
                    const double cj1 = -src[q] - src[q + (size_t)nbw * nq];
                    const double cj2 = dsrc[q] + dsrc[q + nq];
                    const double d2J0 = 2 * cf2[1];
                    const double d2Jb = 6.0 * cf2[4 * nbw - 4] / (nbw * T) + 2.0 * cf2[4 * nbw - 3];
                    const double cj3 = -d2J0 - d2Jb;
                    CJ[which][0] = cj1;
                    CJ[which][1] = cj2;
                    CJ[which][2] = cj3;
                    for (int j = 0; j < nbw / 2; j++) {
                        const double kn = (2 * j + 1) * PI * T;
                        const double ang = (2 * j + 1) * PI / nbw;
                        dst[j + nbw / 2] = -I * cj1 / kn - cj2 / (kn * kn) + I * cj3 / (kn * kn * kn) +
                                           (1.0 - (cos(ang) + I * sin(ang))) * G[j] / (kn * kn * kn * kn);
                    }
                    for (int j = nbw / 2; j < nbw; j++) {
                        const double kn = (2 * (j - nbw) + 1) * PI * T;
                        const double ang = (2 * (j - nbw) + 1) * PI / nbw;
                        dst[j - nbw / 2] = -I * cj1 / kn - cj2 / (kn * kn) + I * cj3 / (kn * kn * kn) +
                                           (1.0 - (cos(ang) + I * sin(ang))) * G[j] / (kn * kn * kn * kn);
                    }
                }
                const double CJ1 = CJ[0][0];
                /* Dyson G, asymptote subtraction (:3349-3407) */
                double complex k0 = 0, qn0 = 0;
                double dGp = 0, dJ0 = 0;
                for (int j = 0; j < nw; j++) {
                    const double kn = (2.0 * (j - nw) + 1) * PI * T;
                    const double complex R = o->Self[ki + (size_t)(nw - j - 1) * nk8];
                    const double complex sg = (-C2 / (kn * kn) + creal(R) / (kn * kn * kn * kn)) +
                                              I * (-o->Sigma_inf / kn + C3 / (kn * kn * kn) - cimag(R) / (kn * kn * kn * kn));
                    const double complex Gv = T / ((I * kn - ekv) + mu - sg);
                    Jt[j] = Gv * Jt[j];
                    qn0 += Gv * Gv * J0[j] / T;
                    const double complex G2 = T / ((-ekv + mu) + I * (kn + o->Sigma_inf / kn));
                    G[j] = Gv - G2;
                    Jt[j] = Jt[j] - G2 * (-I * CJ1 / kn);
                    k0 += G[j];
                    dGp += kn * cimag(G[j]);
                    dJ0 -= kn * cimag(Jt[j]);
                }
                for (int j = nbw - 1; j >= nw; j--) {
                    const double kn = (2.0 * (j - nw) + 1) * PI * T;
                    const double complex R = o->Self[ki + (size_t)(j - nw) * nk8];
                    const double complex sg = (-C2 / (kn * kn) + creal(R) / (kn * kn * kn * kn)) +
                                              I * (-o->Sigma_inf / kn + C3 / (kn * kn * kn) + cimag(R) / (kn * kn * kn * kn));
                    const double complex Gv = T / ((I * kn - ekv) + mu - sg);
                    Jt[j] = Gv * Jt[j];
                    qn0 += Gv * Gv * J0[j] / T;
                    const double complex G2 = T / ((-ekv + mu) + I * (kn + o->Sigma_inf / kn));
                    G[j] = Gv - G2;
                    Jt[j] = Jt[j] - G2 * (-I * CJ1 / kn);
                    k0 += G[j];
                    dGp += kn * cimag(G[j]);
                    dJ0 -= kn * cimag(Jt[j]);
                }
                /* 2-pole asymptotic G and J in tau (:3410-3467) */
                pole_pair pp;
                asym_poles(o, ekv, CJ1, &pp);
                double complex f1p, f1m, fj1, fj1b, f2p, f2m, fj2, fj2b;
                if (creal(pp.r1) > 0) {
                    const double complex e = cexp(-pp.r1 / T);
                    f1p = pp.a1 / pp.den1;
                    f1m = pp.a1 * e / pp.den1;
                    fj1 = pp.r1 * pp.b1 * e / pp.den1;
                    fj1b = pp.r1 * pp.b1 / pp.den1;
                } else {
                    const double complex e = cexp(pp.r1 / T);
                    f1p = pp.a1 * e / pp.den1;
                    f1m = pp.a1 / pp.den1;
                    fj1 = pp.r1 * pp.b1 / pp.den1;
                    fj1b = pp.r1 * pp.b1 * e / pp.den1;
                }
                if (creal(pp.r2) > 0) {
                    const double complex e = cexp(-pp.r2 / T);
                    f2p = pp.a2 / pp.den2;
                    f2m = pp.a2 * e / pp.den2;
                    fj2 = pp.r2 * pp.b2 * e / pp.den2;
                    fj2b = pp.r2 * pp.b2 / pp.den2;
                } else {
                    const double complex e = cexp(pp.r2 / T);
                    f2p = pp.a2 * e / pp.den2;
                    f2m = pp.a2 / pp.den2;
                    fj2 = pp.r2 * pp.b2 / pp.den2;
                    fj2b = pp.r2 * pp.b2 * e / pp.den2;
                }
                k0 += creal(f1m + f2m);
                double dGm = -dGp;
                dGp += creal(pp.r1 * f1p + pp.r2 * f2p);
                dGm += creal(pp.r1 * f1m + pp.r2 * f2m);
                double dJb = -dJ0;
                dJ0 += creal(fj1 + fj2);
                dJb += creal(fj1b + fj2b);
                /* to tau (:3574-3697) */
                orc_dft(G, nbw, -1);
                double complex Gb = G[0];
                orc_dft(Jt, nbw, +1);
                double complex Jb = Jt[0];
                for (int j = 0; j < nbw; j++) {
                    const double t = j / (nbw * T);
                    double complex e1, ej1, e2, ej2;
                    if (creal(pp.r1) > 0) {
                        e1 = cexp(-t * pp.r1);
                        ej1 = cexp(pp.r1 * (t - 1.0 / T));
                    } else {
                        e1 = cexp((1.0 / T - t) * pp.r1);
                        ej1 = cexp(pp.r1 * t);
                    }
                    if (creal(pp.r2) > 0) {
                        e2 = cexp(-t * pp.r2);
                        ej2 = cexp(pp.r2 * (t - 1.0 / T));
                    } else {
                        e2 = cexp((1.0 / T - t) * pp.r2);
                        ej2 = cexp(pp.r2 * t);
                    }
                    const double G0t = creal(-pp.a1 * e1 / pp.den1 - pp.a2 * e2 / pp.den2);
                    const double Jinf = creal(pp.b1 * ej1 / pp.den1 + pp.b2 * ej2 / pp.den2);
                    const double arg = (j * PI * (nbw - 1)) / nbw;
                    const double complex ph = cos(arg) + I * sin(arg);
                    G[j] = ph * G[j] + G0t;
                    Jt[j] = Jt[j] / ph + Jinf;
                }
                {
                    const double t = 1.0 / T;
                    double complex e1, ej1, e2, ej2;
                    if (creal(pp.r1) > 0) {
                        e1 = cexp(-t * pp.r1);
                        ej1 = 1.0;
                    } else {
                        e1 = 1.0;
                        ej1 = cexp(pp.r1 * t);
                    }
                    if (creal(pp.r2) > 0) {
                        e2 = cexp(-t * pp.r2);
                        ej2 = 1.0;
                    } else {
                        e2 = 1.0;
                        ej2 = cexp(pp.r2 * t);
                    }
                    const double G0t = creal(-pp.a1 * e1 / pp.den1 - pp.a2 * e2 / pp.den2);
                    const double Jinf = creal(pp.b1 * ej1 / pp.den1 + pp.b2 * ej2 / pp.den2);
                    const double arg = PI * (nbw - 1);
                    const double complex ph = cos(arg) + I * sin(arg);
                    Gb = ph * Gb + G0t;
                    Jb = Jb / ph + Jinf;
                }
                /* G(tau) J(-tau): spline + DFT -> MT term (:3699-3732) */
                for (int j = 0; j < nbw; j++) {
                    tau2[j] = j / (nbw * T);
                    v2[j] = creal(G[j] * Jt[j]);
                }
                tau2[nbw] = 1.0 / T;
                v2[nbw] = creal((-G[0] - 1.0) * Jb);
                cf2[0] = creal(dGp * Jt[0] + G[0] * dJ0);
                cf2[1] = creal(dGm * Jb - (G[0] + 1.0) * dJb);
                const double MJ1 = v2[0] - v2[nbw];
                orc_spline_coeffs_rel_ws(tau2, v2, nbw + 1, cf2, ws2.B, ws2.r);
                const double d2J0 = 2 * cf2[1];
                const double d2Jb = 6.0 * cf2[4 * nbw - 4] / (nbw * T) + 2.0 * cf2[4 * nbw - 3];
                const double MJ3 = d2J0 - d2Jb;
                for (int j = 0; j < nbw; j++) J0[j] = 6.0 * cf2[4 * j];
                orc_dft(J0, nbw, -1);
                /* -G(tau)G(-tau): spline + DFT -> bubble (:3734-3778) */
                GG[0] = G[0] * (G[0] + 1.0);
                for (int j = 1; j < nbw; j++) GG[j] = -G[j] * G[nbw - j];
                for (int j = 0; j < NS0; j++) {
                    tau[j] = j / (nbw * T);
                    v1[j] = creal(GG[j]);
                }
                const double dchi = dGp * (creal(G[0]) + 1.0) + creal(G[0]) * dGm;
                cf[0] = dchi;
                cf[1] = 0;
                orc_spline_coeffs_rel_ws(tau, v1, NS0, cf, ws1.B, ws1.r);
                for (int j = 0; j < nbw / 2; j++) G[j] = 6.0 * cf[4 * j];
                for (int j = nbw / 2; j < nbw; j++) G[j] = -G[nbw - j - 1];
                orc_dft(G, nbw, -1);
                double integ = 0;
                for (int j = 0; j < nbw / 2; j++) {
                    const double a = cf[4 * j], bb = cf[4 * j + 1], c = cf[4 * j + 2], d = cf[4 * j + 3];
                    const double x1 = tau[j + 1] - tau[j];
                    integ += a * x1 * x1 * x1 * x1 / 4.0 + bb * x1 * x1 * x1 / 3.0 + c * x1 * x1 / 2.0 + d * x1;
                }
                /* accumulate (:3780-3822) */
                double wt = 4;
                if (m == 0 || m == nbk - 1) wt = wt / 2;
                k0acc -= 2.0 * wt * normFact * d2ek[m + (l + 1) * nbk] * creal(k0);
                const double bub0 = 2 * integ;
                c1[0] -= 2.0 * wt * normFact * v * v * bub0;
                inf1 += 4.0 * dchi * normFact * wt * v * v;
                inf2 += normFact * wt * v * 2.0 * MJ1;
                inf4 -= normFact * wt * v * 2.0 * MJ3;
                c2[0] -= normFact * wt * v * creal(qn0);
                const double th = atan((1.0 * (nk0 / 2 - m)) / (nk0 / 2 - l - 1));
                int it = (int)floor(th / Dtheta);
                if (it >= N_theta) it = N_theta - 1;
                if (it < 0) it = 0;
                c1k[it] -= 2.0 * wt * normFact * v * v * bub0;
                c2k[it] -= normFact * wt * v * creal(qn0);
                for (int j = 1; j <= nw; j++) {
                    const double wn = 2 * j * PI * T;
                    const double ang = -2 * j * PI / nbw;
                    const double complex e = 1.0 - (cos(ang) + I * sin(ang));
                    const double imJ = -2.0 * MJ1 / wn + 2.0 * MJ3 / (wn * wn * wn) + 2.0 * cimag(e * J0[j] / (wn * wn * wn * wn));
                    c2[j] -= normFact * wt * v * imJ / wn;
                    c2k[it + (size_t)j * N_theta] -= normFact * wt * v * imJ / wn;
                    const double bub = -2.0 * dchi / (wn * wn) + creal(e * G[j] / (wn * wn * wn * wn));
                    c1[j] -= 2.0 * normFact * wt * v * v * bub;
                    c1k[it + (size_t)j * N_theta] -= 2.0 * normFact * wt * v * v * bub;
                }
            }
            pk0[b] = k0acc;
            pinf[b][0] = inf1;
            pinf[b][1] = inf2;
            pinf[b][2] = inf4;
        }
        free(Jt);
        free(J0);
        free(G);
        free(GG);
        free(tau2);
        free(v2);
        free(cf2);
        free(tau);
        free(v1);
        free(cf);
        sws_free(&ws2);
        sws_free(&ws1);
    }
    free(o->chijj1);
    free(o->chijj2);
    free(o->chijj1_k);
    free(o->chijj2_k);
    o->chijj1 = dcalloc(nwb);
    o->chijj2 = dcalloc(nwb);
    o->chijj1_k = dcalloc((size_t)nwb * N_theta);
    o->chijj2_k = dcalloc((size_t)nwb * N_theta);
    double k0sum = 0;
    o->chijj1_inf2 = o->chijj2_inf2 = o->chijj2_inf4 = 0;
    for (int b = 0; b < NBLK; b++) {
        for (int j = 0; j < nwb; j++) {
            o->chijj1[j] += p1[(size_t)b * nwb + j];
            o->chijj2[j] += p2[(size_t)b * nwb + j];
        }
        for (size_t j = 0; j < (size_t)nwb * N_theta; j++) {
            o->chijj1_k[j] += p1k[(size_t)b * nwb * N_theta + j];
            o->chijj2_k[j] += p2k[(size_t)b * nwb * N_theta + j];
        }
        k0sum += pk0[b];
        o->chijj1_inf2 += pinf[b][0];
        o->chijj2_inf2 += pinf[b][1];
        o->chijj2_inf4 += pinf[b][2];
    }
    free(p1);
    free(p2);
    free(p1k);
    free(p2k);

    /* kx = 0 and kx = pi rows of the diamagnetic term (cond_opt.cpp:3993-4173) */
    for (int l = 0; l < nbk; l += nk0 / 2) {
        for (int m = 0; m < nbk; m++) {
            const long ki = fold_idx(l, m, nk0);
            const double C2 = o->Sigma_inf2[ki], C3 = o->Sigma_inf3[ki];
            const double ekv = o->ek[m + l * nbk];
            double complex k0 = 0;
            for (int j = 0; j < nw; j++) {
                const double kn = (2.0 * (j - nw) + 1) * PI * T;
                const double complex R = o->Self[ki + (size_t)(nw - j - 1) * nk8];
                const double complex sg = (-C2 / (kn * kn) + creal(R) / (kn * kn * kn * kn)) +
                                          I * (-o->Sigma_inf / kn + C3 / (kn * kn * kn) - cimag(R) / (kn * kn * kn * kn));
                const double complex Gv = T / ((I * kn - ekv) + mu - sg);
                const double complex G0v = T / ((-ekv + mu) + I * (kn + o->Sigma_inf / kn));
                k0 += Gv - G0v;
            }
            for (int j = nbw - 1; j >= nw; j--) {
                const double kn = (2.0 * (j - nw) + 1) * PI * T;
                const double complex R = o->Self[ki + (size_t)(j - nw) * nk8];
                const double complex sg = (-C2 / (kn * kn) + creal(R) / (kn * kn * kn * kn)) +
                                          I * (-o->Sigma_inf / kn + C3 / (kn * kn * kn) + cimag(R) / (kn * kn * kn * kn));
                const double complex Gv = T / ((I * kn - ekv) + mu - sg);
                const double complex G0v = T / ((-ekv + mu) + I * (kn + o->Sigma_inf / kn));
                k0 += Gv - G0v;
            }
            pole_pair pp;
            asym_poles(o, ekv, 0.0, &pp);
            const double complex f1m = creal(pp.r1) > 0 ? pp.a1 * cexp(-pp.r1 / T) / pp.den1 : pp.a1 / pp.den1;
            const double complex f2m = creal(pp.r2) > 0 ? pp.a2 * cexp(-pp.r2 / T) / pp.den2 : pp.a2 / pp.den2;
            k0 += creal(f1m + f2m);
            double wt = 2;
            if (m == 0 || m == nbk - 1) wt = wt / 2;
            k0sum -= 2.0 * wt * normFact * d2ek[m + l * nbk] * creal(k0);
        }
    }
    o->kx0 = k0sum;
    free(h);
    free(f0);
    free(dhV);
    free(df0V);
    free(dek);
    free(d2ek);
    free(ekmu);
    return 0;
}

/* ---------------------------------------------------------------- full solve */
/* tpsc_main.cpp:61-75 (green ctor, calc_Self_FFT_spline, traceSelfG, find_mu) then the
   Lambda_jj call of tpsc_main.cpp:243 (chi3 = false, N_theta_k = 1) */
int orc_solve(orc *o)
{
    int ok = orc_find_Usp_Uch(o);
    orc_vertex(o);
    orc_self_energy(o);
    orc_trace_self_G(o);
    ok &= orc_find_mu(o);
    orc_chijj(o, o->N_theta > 0 ? o->N_theta : 1);
    return ok;
}

/* orc_solve without the coarse pass, with the wall time of its stages in t[6] (seconds):
   A mu0 + chi0, B Usp/Uch roots, C vertex, D self-energy, E traceSelfG + mu, F Lambda_jj */
int orc_solve_timed(orc *o, double t[6])
{
    double w0 = omp_get_wtime();
    orc_find_mu0(o);
    orc_chi0(o, o->p.nk0, o->p.Nw0);
    double w1 = omp_get_wtime();
    t[0] = w1 - w0;
    int ok = orc_find_Usp_Uch_impl(o, 1);
    w0 = omp_get_wtime();
    t[1] = w0 - w1;
    orc_vertex(o);
    w1 = omp_get_wtime();
    t[2] = w1 - w0;
    orc_self_energy(o);
    w0 = omp_get_wtime();
    t[3] = w0 - w1;
    orc_trace_self_G(o);
    ok &= orc_find_mu(o);
    w1 = omp_get_wtime();
    t[4] = w1 - w0;
    orc_chijj(o, o->N_theta > 0 ? o->N_theta : 1);
    t[5] = omp_get_wtime() - w1;
    return ok;
}

/* The chain after the sum rules from given roots, on the chi0 of a solve already done: Usp as
   given, Uch as given or, if NaN, by the charge-rule search (chi.cpp:866-869), the double
   occupancy (chi.cpp:370-371), then vertex, self-energy, traceSelfG, mu and Lambda_jj as
   orc_solve runs them.  For checking a solve whose roots are not the ones orc_find_Usp_Uch
   returns, such as where that search stops short of a root. */
int orc_solve_from_roots(orc *o, double Usp, double Uch)
{
    o->Usp = Usp;
    int ok = 1;
    if (isnan(Uch)) ok = orc_uch_root(o, &Uch);
    o->Uch = Uch;
    const double U = o->p.U, nsr = o->nsr;
    if (o->p.n > 1.0)
        o->dblOcc = o->Usp / U * nsr * nsr / 4.0 + 1.0 - nsr;
    else
        o->dblOcc = o->Usp / U * nsr * nsr / 4.0;
    orc_vertex(o);
    orc_self_energy(o);
    orc_trace_self_G(o);
    ok &= orc_find_mu(o);
    orc_chijj(o, o->N_theta > 0 ? o->N_theta : 1);
    return ok;
}

int orc_omp_max_threads(void) { return omp_get_max_threads(); }

void orc_self_kw_brute_out(orc *o, int kx, int ky, int n, double out[2])
{
    const double complex s = orc_self_kw_brute(o, kx, ky, n);
    out[0] = creal(s);
    out[1] = cimag(s);
}

/* ================================================================ a11: Aslamazov-Larkin */
/* Sparse boson set of the AL term (cond_opt.cpp:5123-5156): n = 0 .. (N0-1)/2^m, then m
   groups of (N0-1)/2^(m+1) frequencies with steps 2, 4, ..., 2^m; N0 = Nw/2 + 1.
   Returns the count; writes the indices to out when out != NULL. */
int orc_al_freqs(int nbw, int mmin, long *out)
{
    const int N0 = nbw / 2 + 1;
    const long pw = 1L << mmin;
    const int jf0 = (int)((N0 - 1) / pw);
    const int Nj = (int)((N0 - 1) / (2 * pw));
    const int N = mmin * Nj + jf0 + 1;
    if (out) {
        for (int j = 0; j <= jf0; j++) out[j] = j;
        long n = jf0, d = 1;
        int jf = jf0;
        for (int l = 0; l < mmin; l++) {
            const int j0 = jf + 1;
            jf = j0 + Nj - 1;
            d *= 2;
            for (int j = j0; j <= jf; j++) {
                n += d;
                out[j] = n;
            }
        }
    }
    return N;
}

/* chi0(q,iq_p) with the 1/qn^2 + 1/qn^4 tail beyond the grid (cond_opt.cpp:5383-5392) */
static inline double al_chi0(const orc *o, long r, long p, long nk8, int H, double T)
{
    if (p < 0) p = -p;
    if (p <= H) return o->chiqw[r + p * nk8];
    const double wn = 2 * PI * p * T;
    return o->chi0_inf[r] / (wn * wn) + o->chi0_inf2[r] / (wn * wn * wn * wn);
}
/* hG(q,iq_p) from its transform samples and 1/qn^2 coefficient (cond_opt.cpp:5396-5401) */
static inline double al_hG(const double *hw, const double *hg2, long q, long p, long nq, int H, double T)
{
    if (p < 0) p = -p;
    if (p == 0) return hw[q];
    const double wn = 2 * PI * p * T;
    const double c4 = (p < H) ? hw[q + p * nq] : hw[q + (long)(H - 1) * nq];
    return -hg2[q] / (wn * wn) + c4 / (wn * wn * wn * wn);
}
/* Sigma(k,iw_f) from the moments and the remainder, conjugated for f < 0
   (cond_opt.cpp:6314-6377) */
static inline double complex al_sigma(const orc *o, long ki, long f, long nk8, int H)
{
    const double kn = (2 * f + 1) * PI * o->p.T;
    const double C2 = o->Sigma_inf2[ki], C3 = o->Sigma_inf3[ki];
    double re = -C2 / (kn * kn), im = -o->Sigma_inf / kn + C3 / (kn * kn * kn);
    const long g = (f >= 0) ? f : -f - 1;
    if (g < H) {
        const double complex R = o->Self[ki + g * nk8];
        re += creal(R) / (kn * kn * kn * kn);
        im += ((f >= 0) ? cimag(R) : -cimag(R)) / (kn * kn * kn * kn);
    }
    return re + I * im;
}

/* Imaginary-time form of the two-pole tail Hinf(iq_p)/hg2 = 3 Usp/(-w^2 + a^2) + Uch/(-w^2 - c^2),
   a^2 = Usp chi0_inf/2, c^2 = Uch chi0_inf/2, i.e. T sum_p e^{-i w_p tau} Hinf(iq_p)/hg2, in the
   pole-sum form of cond_opt.cpp:5556-5618 (poles zsp = +-i a, zch = +-c).  kind 0: value at
   0 <= tau < beta (:5562-5575); 1: value at tau = beta (:5577-5590); 2: slope at tau = 0+
   (:5592-5604); 3: slope at tau = beta- (:5606-5618). */
double complex orc_al_tail(double Usp, double Uch, double chi0inf, double T, double tau, int kind)
{
    const double rt = sqrt(Usp * chi0inf / 2.0);
    const double complex zs0 = I * rt, zs1 = -zs0;
    const double zc0 = sqrt(Uch * chi0inf / 2.0), zc1 = -zc0;
    const double b = 1.0 / T;
    double complex tsp;
    double tch;
    if (kind == 0) {
        tsp = cexp(-zs0 * tau) / ((cexp(-zs0 * b) - 1.0) * (zs0 - zs1)) + cexp(-zs1 * tau) / ((cexp(-zs1 * b) - 1.0) * (zs1 - zs0));
        tch = exp(-zc0 * tau) / ((exp(-zc0 * b) - 1.0) * (zc0 - zc1)) + exp(zc1 * (b - tau)) / ((1.0 - exp(zc1 * b)) * (zc1 - zc0));
    } else if (kind == 1) {
        const double complex e0 = cexp(zs0 * b), e1 = cexp(zs1 * b);
        tsp = 1.0 / ((1.0 - e0) * (zs0 - zs1)) + 1.0 / ((1.0 - e1) * (zs1 - zs0));
        const double ec = exp(-zc0 * b);
        tch = ec / ((ec - 1.0) * (zc0 - zc1)) + 1.0 / ((1.0 - exp(zc1 * b)) * (zc1 - zc0));
    } else if (kind == 2) {
        tsp = -zs0 / ((cexp(-zs0 * b) - 1.0) * (zs0 - zs1)) - zs1 / ((cexp(-zs1 * b) - 1.0) * (zs1 - zs0));
        const double e1 = exp(zc1 * b);
        tch = -zc0 / ((exp(-zc0 * b) - 1.0) * (zc0 - zc1)) - zc1 * e1 / ((1.0 - e1) * (zc1 - zc0));
    } else {
        const double complex e0 = cexp(-zs0 * b), e1 = cexp(-zs1 * b);
        tsp = -zs0 * e0 / ((e0 - 1.0) * (zs0 - zs1)) - zs1 * e1 / ((e1 - 1.0) * (zs1 - zs0));
        const double ec = exp(-zc0 * b);
        tch = -zc0 * ec / ((ec - 1.0) * (zc0 - zc1)) - zc1 / ((1.0 - exp(zc1 * b)) * (zc1 - zc0));
    }
    return 3.0 * Usp * tsp + Uch * tch;
}
void orc_al_tail_out(double Usp, double Uch, double chi0inf, double T, double tau, int kind, double out[2])
{
    const double complex v = orc_al_tail(Usp, Uch, chi0inf, T, tau, kind);
    out[0] = creal(v);
    out[1] = cimag(v);
}

/* spline + shifted fermionic transform of a complex tau-pencil z_j, j = 0..nbw, with
   boundary slopes d0, dN (cond_opt.cpp:5907-5995 for HG, :6172-6245 for HG2):
   out[j] = value at w_(f(j)+shift), f(j) = j (j < nbw/2 - nc) or j - nbw */
static void al_k_pencil(const double complex *z, double complex d0, double complex dN, int nbw, double T,
                        long shift, int nc, double *tau, double *v, double *cf, double complex *ft, sws *ws,
                        double complex *out)
{
    const int H = nbw / 2;
    double complex d2S0 = 0, d2Sb = 0;
    for (int part = 0; part < 2; part++) {
        for (int j = 0; j <= nbw; j++) {
            tau[j] = j / (nbw * T);
            v[j] = part ? cimag(z[j]) : creal(z[j]);
        }
        cf[0] = part ? cimag(d0) : creal(d0);
        cf[1] = part ? cimag(dN) : creal(dN);
        orc_spline_coeffs_rel_ws(tau, v, nbw + 1, cf, ws->B, ws->r);
        const double s0 = 2 * cf[1];
        const double sb = 6.0 * cf[4 * nbw - 4] / (nbw * T) + 2.0 * cf[4 * nbw - 3];
        if (!part) {
            d2S0 = s0;
            d2Sb = sb;
            for (int j = 0; j < nbw; j++) ft[j] = cexp(-I * PI * j / nbw) * (6.0 * cf[4 * j]);
        } else {
            d2S0 += I * s0;
            d2Sb += I * sb;
            for (int j = 0; j < nbw; j++) ft[j] += I * cexp(-I * PI * j / nbw) * (6.0 * cf[4 * j]);
        }
    }
    orc_dft(ft, nbw, -1);
    const double complex cw = z[0] + z[nbw], ds = d0 + dN, d2 = d2S0 + d2Sb;
    for (int j = 0; j < nbw; j++) {
        const long f = (j < H - nc) ? j : j - nbw;
        const long s = f + shift;
        const double wn = (2 * s + 1) * PI * T;
        double complex val = -I * cw / wn - ds / (wn * wn) + I * d2 / (wn * wn * wn);
        if (s < nbw) {
            const long fi = (s < 0) ? s + nbw : s;
            val += (1.0 - cexp(-I * (2 * (j + shift) + 1) * PI / nbw)) * ft[fi] / (wn * wn * wn * wn);
        }
        out[j] = val;
    }
}

/* cond_opt.cpp:4292-6467 with chi3 = true: the Aslamazov-Larkin term Lambda_jj3 on the
   sparse boson set, after orc_chijj.  Per-k contributions are summed in k order. */
int orc_chijj_al(orc *o, int mmin, int N_theta)
{
    if (!o->chijj1 || !o->Self || !o->chiqw || mmin < 0 || mmin > 24 || N_theta < 1) return -1;
    const int nbk = o->nbk, nbw = o->nbw, H = nbw / 2, nk0 = 2 * (nbk - 1);
    const long nq = (long)nbk * (nbk - 2), nk8 = (long)nbk * (nbk + 1) / 2;
    const double T = o->p.T, mu0 = o->mu0, mu = o->mu, U = o->p.U, Usp = o->Usp, Uch = o->Uch;
    const double nf = 1.0 / ((double)nk0 * nk0);
    const double Dtheta = PI / (2.0 * N_theta);
    const int Ns = orc_al_freqs(nbw, mmin, NULL);

    free(o->al_freq);
    free(o->chijj3);
    free(o->chijj_all);
    free(o->chijj3_k);
    free(o->al_hw);
    free(o->al_hg2);
    o->al_freq = (long *)xmalloc(sizeof(long) * (size_t)Ns);
    orc_al_freqs(nbw, mmin, o->al_freq);
    o->al_N = Ns;
    o->al_N_theta = N_theta;
    o->chijj3 = dcalloc((size_t)Ns);
    o->chijj_all = dcalloc((size_t)Ns);
    o->chijj3_k = dcalloc((size_t)Ns * N_theta);

    double *dek = dalloc((size_t)nq);
    for (int l = 0; l < nbk - 2; l++)
        for (int m = 0; m < nbk; m++) dek[m + l * nbk] = orc_grad_ek_x(o, (l + 1) * PI / (nbk - 1), m * PI / (nbk - 1));

    /* h(r,tau) (cond_opt.cpp:4302-4330) and its slopes dh (:2745-2781) */
    double *h = dalloc((size_t)nq * (nbw + 1));
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= nbw; j++) {
        const double t = (double)j / (T * nbw);
        double *fh = h + (size_t)j * nq;
        for (int l = 0; l < nbk - 2; l++)
            for (int m = 0; m < nbk; m++) {
                const double e = o->ek[m + (l + 1) * nbk], v = dek[m + l * nbk];
                if (e > mu0)
                    fh[m + l * nbk] = nf * v * exp((mu0 - e) * t) / (exp((mu0 - e) / T) + 1.0);
                else
                    fh[m + l * nbk] = nf * v * exp((e - mu0) * (1.0 / T - t)) / (exp((e - mu0) / T) + 1.0);
            }
        orc_r2r_2d(fh, nbk - 2, nbk, 1, 0);
    }
    double *dh = dalloc(2 * (size_t)nq);
    for (int l = 0; l < nbk - 2; l++)
        for (int m = 0; m < nbk; m++) {
            const long q = m + l * nbk;
            const double e = o->ek[m + (l + 1) * nbk], v = dek[q];
            if (e > mu0) {
                const double ex = exp((mu0 - e) / T);
                dh[q] = -nf * v * (e - mu0) / (ex + 1.0);
                dh[q + nq] = -nf * v * (e - mu0) * ex / (ex + 1.0);
            } else {
                const double ex = exp((e - mu0) / T);
                dh[q] = -nf * v * (e - mu0) * ex / (ex + 1.0);
                dh[q + nq] = -nf * v * (e - mu0) / (ex + 1.0);
            }
        }
    orc_r2r_2d(dh, nbk - 2, nbk, 1, 0);
    orc_r2r_2d(dh + nq, nbk - 2, nbk, 1, 0);

    /* G0(r,-tau) on the wedge (cond_opt.cpp:4374-4413) and its slopes (:4415-4445) */
    double *Grt = dalloc((size_t)nk8 * (nbw + 1)), *dGrt = dalloc(2 * (size_t)nk8);
#pragma omp parallel
    {
        double *sq = dalloc((size_t)nbk * nbk);
#pragma omp for schedule(static)
        for (int j = 0; j <= nbw; j++) {
            const double t = (double)j / (T * nbw);
            for (int l = 0; l < nbk; l++)
                for (int m = 0; m <= l; m++) {
                    const double x = o->ek[m + l * nbk] - mu0;
                    const double g = (x < 0) ? nf * exp(x * t) / (exp(x / T) + 1.0)
                                             : nf * exp(x * (t - 1.0 / T)) / (exp(-x / T) + 1.0);
                    sq[m + l * nbk] = g;
                    sq[l + m * nbk] = g;
                }
            orc_r2r_2d(sq, nbk, nbk, 0, 0);
            for (int l = 0; l < nbk; l++)
                for (int m = 0; m <= l; m++) Grt[widx(l, m) + (size_t)j * nk8] = sq[m + l * nbk];
        }
        free(sq);
    }
    {
        double *sq = dalloc((size_t)nbk * nbk);
        for (int l = 0; l < nbk; l++)
            for (int m = 0; m <= l; m++) {
                const double x = o->ek[m + l * nbk] - mu0;
                const double g = (x < 0) ? nf * x / (exp(x / T) + 1.0) : nf * x * exp(-x / T) / (exp(-x / T) + 1.0);
                sq[m + l * nbk] = g;
                sq[l + m * nbk] = g;
            }
        orc_r2r_2d(sq, nbk, nbk, 0, 0);
        for (int l = 0; l < nbk; l++)
            for (int m = 0; m <= l; m++) dGrt[widx(l, m)] = sq[m + l * nbk];
        for (long w = 0; w < nk8; w++) dGrt[nk8 + w] = ((w == 0) ? -mu0 : o->er[w]) - dGrt[w];
        free(sq);
    }

    /* slopes of h(r,tau) G(r,-tau) -> q (cond_opt.cpp:5159-5229), then hG(q,tau) (:5231-5261) */
    double *dhG = dalloc(2 * (size_t)nq);
    for (int side = 0; side < 2; side++) {
        const int jj = side ? nbw : 0;
        double *d = dhG + side * nq;
        for (int l = 0; l < nbk - 2; l++)
            for (int m = 0; m < nbk; m++) {
                const long q = m + l * nbk, r = fold_idx(l + 1, m, nk0);
                d[q] = dh[q + side * nq] * Grt[r + (size_t)jj * nk8] + h[q + (size_t)jj * nq] * dGrt[r + side * nk8];
            }
        orc_r2r_2d(d, nbk - 2, nbk, 1, 0);
    }
#pragma omp parallel for schedule(static)
    for (int j = 0; j <= nbw; j++) {
        double *fh = h + (size_t)j * nq;
        for (int l = 0; l < nbk - 2; l++)
            for (int m = 0; m < nbk; m++) fh[m + l * nbk] *= Grt[fold_idx(l + 1, m, nk0) + (size_t)j * nk8];
        orc_r2r_2d(fh, nbk - 2, nbk, 1, 0);
    }

    /* hG(q,iq_p), p < nbw/2, and its 1/qn^2 coefficient hg2 (cond_opt.cpp:5266-5330) */
    double *hw = dalloc((size_t)nq * H), *hg2 = dalloc((size_t)nq);
#pragma omp parallel
    {
        double *tau = dalloc((size_t)nbw + 1), *v = dalloc((size_t)nbw + 1), *cf = dalloc(4 * (size_t)nbw);
        double complex *ft = (double complex *)xmalloc(sizeof(double complex) * (size_t)nbw);
        sws ws = sws_new(nbw + 1);
#pragma omp for schedule(static)
        for (long q = 0; q < nq; q++) {
            for (int j = 0; j <= nbw; j++) {
                tau[j] = j / (nbw * T);
                v[j] = h[q + (size_t)j * nq];
            }
            cf[0] = dhG[q];
            cf[1] = dhG[q + nq];
            orc_spline_coeffs_rel_ws(tau, v, nbw + 1, cf, ws.B, ws.r);
            for (int j = 0; j < nbw; j++) ft[j] = 6.0 * cf[4 * j];
            orc_dft(ft, nbw, -1);
            double integ = 0;
            for (int j = 0; j < nbw; j++) {
                const double a = cf[4 * j], b = cf[4 * j + 1], c = cf[4 * j + 2], d = cf[4 * j + 3];
                const double x1 = tau[j + 1] - tau[j];
                integ += a * x1 * x1 * x1 * x1 / 4.0 + b * x1 * x1 * x1 / 3.0 + c * x1 * x1 / 2.0 + d * x1;
            }
            hw[q] = 2 * integ;
            hg2[q] = 2 * (dhG[q] - dhG[q + nq]);
            for (int j = 1; j < H; j++) hw[q + (size_t)j * nq] = 2 * creal((1.0 - cexp(-I * 2.0 * j * PI / nbw)) * ft[j]);
        }
        free(tau);
        free(v);
        free(cf);
        free(ft);
        sws_free(&ws);
    }
    o->al_hw = hw;
    o->al_hg2 = hg2;
    free(dh);

    /* per boson frequency of the sparse set (cond_opt.cpp:5350-6400) */
    const size_t tsz = (size_t)nq * (nbw + 1);
    double complex *HG = (double complex *)xmalloc(sizeof(double complex) * tsz);
    double complex *HR = (double complex *)xmalloc(sizeof(double complex) * tsz);  /* HG(r,tau) */
    double complex *HK = (double complex *)xmalloc(sizeof(double complex) * tsz);  /* products, k space */
    double complex *HK2 = (double complex *)xmalloc(sizeof(double complex) * tsz);
    double complex *dHn = (double complex *)xmalloc(sizeof(double complex) * 2 * nq);
    double complex *dHr = (double complex *)xmalloc(sizeof(double complex) * 2 * nq);
    double complex *dK = (double complex *)xmalloc(sizeof(double complex) * 2 * nq);
    double complex *dK2 = (double complex *)xmalloc(sizeof(double complex) * 2 * nq);
    double complex *contrib = (double complex *)xmalloc(sizeof(double complex) * nq);
    for (int n = 1; n < Ns; n++) {
        const long nn = o->al_freq[n];
        const double qn = 2.0 * nn * PI * T;
        const int nc = (int)(nn / 2);
        /* (a) HG(q,iq_p) on the window p = j (j < nbw/2 - nc), j - nbw (:5350-5490) and (b) its
           transform to tau with the two-pole tails (:5492-5620) */
#pragma omp parallel
        {
            double complex *ft = (double complex *)xmalloc(sizeof(double complex) * (size_t)nbw);
            double *Dv = dalloc((size_t)nbw);
#pragma omp for schedule(static)
            for (long q = 0; q < nq; q++) {
                const int l = (int)(q / nbk), m = (int)(q % nbk);
                const long r = fold_idx(l + 1, m, nk0);
                double complex *Hq = HG + q;  /* stride nq */
                for (int j = 0; j < nbw; j++) {
                    const long p1 = (j < H - nc) ? j : j - nbw, p2 = p1 + nn;
                    const double c1 = al_chi0(o, r, p1, nk8, H, T), c2 = al_chi0(o, r, p2, nk8, H, T);
                    const double g1 = al_hG(hw, hg2, q, p1, nq, H, T), g2 = al_hG(hw, hg2, q, p2, nq, H, T);
                    Dv[j] = (g1 - g2) / (c1 - c2);
                }
                for (int j = 0; j < nbw; j++) {
                    const long p1 = (j < H - nc) ? j : j - nbw;
                    const double c1 = al_chi0(o, r, p1, nk8, H, T);
                    const double chisp = c1 / (1.0 - Usp * c1 / 2), chich = c1 / (1.0 + Uch * c1 / 2);
                    const double G = 3 * Usp * chisp + Uch * chich;
                    if (nn % 2 == 0 && p1 == -nn / 2)
                        Hq[(size_t)j * nq] = G * (4 * Dv[j - 1] - Dv[j - 2]) / 3.0;
                    else
                        Hq[(size_t)j * nq] = G * Dv[j];
                }
                /* (b) */
                const double cq = -o->chi0_inf[r];
                double complex dtauH0 = 0;
                for (int j = 0; j < nbw; j++) {
                    const long p1 = (j < H - nc) ? j : j - nbw;
                    const double wn = 2 * p1 * PI * T;
                    const double csp = 1.0 / (-wn * wn - Usp * cq / 2.0), cch = 1.0 / (-wn * wn + Uch * cq / 2.0);
                    const double Hinf = hg2[q] * (3.0 * Usp * csp + Uch * cch);
                    ft[j] = T * (Hq[(size_t)j * nq] - Hinf);
                    dtauH0 -= wn * ft[j];
                }
                orc_dft(ft, nbw, -1);
                const double ci = o->chi0_inf[r], g2q = hg2[q];
                for (int j = 0; j < nbw; j++)
                    Hq[(size_t)j * nq] = ft[j] + g2q * orc_al_tail(Usp, Uch, ci, T, j / (nbw * T), 0);
                Hq[(size_t)nbw * nq] = ft[0] + g2q * orc_al_tail(Usp, Uch, ci, T, 1.0 / T, 1);
                dHn[q] = I * dtauH0 + g2q * orc_al_tail(Usp, Uch, ci, T, 0.0, 2);
                dHn[q + nq] = I * dtauH0 + g2q * orc_al_tail(Usp, Uch, ci, T, 1.0 / T, 3);
            }
            free(ft);
            free(Dv);
        }
        /* (c) q -> r of every tau slice and of the slopes (:5622-5720) */
#pragma omp parallel
        {
            double *sr = dalloc((size_t)nq), *si = dalloc((size_t)nq);
#pragma omp for schedule(static)
            for (int j = 0; j <= nbw + 2; j++) {
                const double complex *src = (j <= nbw) ? HG + (size_t)j * nq : dHn + (size_t)(j - nbw - 1) * nq;
                double complex *dst = (j <= nbw) ? HR + (size_t)j * nq : dHr + (size_t)(j - nbw - 1) * nq;
                for (long q = 0; q < nq; q++) {
                    sr[q] = nf * creal(src[q]);
                    si[q] = nf * cimag(src[q]);
                }
                orc_r2r_2d(sr, nbk - 2, nbk, 1, 0);
                orc_r2r_2d(si, nbk - 2, nbk, 1, 0);
                for (long q = 0; q < nq; q++) dst[q] = sr[q] + I * si[q];
            }
            free(sr);
            free(si);
        }
        /* (d)-(f) slopes of HG(r,tau) G(r,-tau) and of HG2(r,tau) G(r,-tau), HG2(r,tau) =
           HG(r,beta-tau), then the products back to k (:5722-5905, :6000-6170) */
#pragma omp parallel
        {
            double *sr = dalloc((size_t)nq), *si = dalloc((size_t)nq);
#pragma omp for schedule(static)
            for (int j = 0; j <= 2 * nbw + 1 + 4; j++) {
                const double complex *A, *B;
                const double *ga, *gb;
                double complex *dst;
                double sgn = 1.0;
                if (j <= nbw) { /* HG(r,tau_j) G(r,-tau_j) */
                    A = HR + (size_t)j * nq; ga = Grt + (size_t)j * nk8; B = NULL; gb = NULL;
                    dst = HK + (size_t)j * nq;
                } else if (j <= 2 * nbw + 1) { /* HG2(r,tau_i) G(r,-tau_i) */
                    const int i = j - nbw - 1;
                    A = HR + (size_t)(nbw - i) * nq; ga = Grt + (size_t)i * nk8; B = NULL; gb = NULL;
                    dst = HK2 + (size_t)i * nq;
                } else { /* slopes: c = 0: HG side 0, 1: HG side beta, 2: HG2 side 0, 3: HG2 side beta */
                    const int c = j - 2 * nbw - 2, side = c & 1;
                    const int which = c >> 1;
                    const int jt = which ? (side ? 0 : nbw) : (side ? nbw : 0);  /* HG time index */
                    A = HR + (size_t)jt * nq;
                    ga = dGrt + (size_t)side * nk8;
                    B = which ? dHr + (size_t)(1 - side) * nq : dHr + (size_t)side * nq;
                    sgn = which ? -1.0 : 1.0;
                    gb = Grt + (size_t)(side ? nbw : 0) * nk8;
                    dst = (which ? dK2 : dK) + (size_t)side * nq;
                }
                for (int l = 0; l < nbk - 2; l++)
                    for (int m = 0; m < nbk; m++) {
                        const long q = m + l * nbk, r = fold_idx(l + 1, m, nk0);
                        double vr = creal(A[q]) * ga[r], vi = cimag(A[q]) * ga[r];
                        if (B) {
                            vr += sgn * creal(B[q]) * gb[r];
                            vi += sgn * cimag(B[q]) * gb[r];
                        }
                        sr[q] = vr;
                        si[q] = vi;
                    }
                orc_r2r_2d(sr, nbk - 2, nbk, 1, 0);
                orc_r2r_2d(si, nbk - 2, nbk, 1, 0);
                for (long q = 0; q < nq; q++) dst[q] = sr[q] + I * si[q];
            }
            free(sr);
            free(si);
        }
        /* (g) per k: HG(k, iw_(f+n)), HG2(k, iw_f) by splines and (h) the sum over the
           fermionic window with G(k,iw_f) G(k,iw_(f+n)) (:5907-6400) */
#pragma omp parallel
        {
            double *tau = dalloc((size_t)nbw + 1), *v = dalloc((size_t)nbw + 1), *cf = dalloc(4 * (size_t)nbw);
            double complex *ft = (double complex *)xmalloc(sizeof(double complex) * (size_t)nbw);
            double complex *z = (double complex *)xmalloc(sizeof(double complex) * ((size_t)nbw + 1));
            double complex *A = (double complex *)xmalloc(sizeof(double complex) * (size_t)nbw);
            double complex *B = (double complex *)xmalloc(sizeof(double complex) * (size_t)nbw);
            sws ws = sws_new(nbw + 1);
#pragma omp for schedule(static)
            for (long q = 0; q < nq; q++) {
                const int l = (int)(q / nbk), m = (int)(q % nbk);
                const long ki = fold_idx(l + 1, m, nk0);
                for (int j = 0; j <= nbw; j++) z[j] = HK[q + (size_t)j * nq];
                al_k_pencil(z, dK[q], dK[q + nq], nbw, T, nn, nc, tau, v, cf, ft, &ws, A);
                for (int j = 0; j <= nbw; j++) z[j] = HK2[q + (size_t)j * nq];
                al_k_pencil(z, dK2[q], dK2[q + nq], nbw, T, 0, nc, tau, v, cf, ft, &ws, B);
                const double e = o->ek[m + (l + 1) * nbk];
                const double wt = (m == 0 || m == nbk - 1) ? 2.0 : 4.0;
                double complex acc = 0;
                for (int j = 0; j < nbw; j++) {
                    const long f = (j < H - nc) ? j : j - nbw;
                    const double k1 = (2 * f + 1) * PI * T, k2 = (2 * (f + nn) + 1) * PI * T;
                    const double complex G1 = 1.0 / (I * k1 - e + mu - al_sigma(o, ki, f, nk8, H));
                    const double complex G2 = 1.0 / (I * k2 - e + mu - al_sigma(o, ki, f + nn, nk8, H));
                    acc -= I * U * T * nf * wt * dek[q] * G1 * G2 * (A[j] - B[j]) / (2 * qn);
                }
                contrib[q] = acc;
            }
            free(tau);
            free(v);
            free(cf);
            free(ft);
            free(z);
            free(A);
            free(B);
            sws_free(&ws);
        }
        double complex s = 0;
        for (long q = 0; q < nq; q++) {
            s += contrib[q];
            const int l = (int)(q / nbk), m = (int)(q % nbk);
            const double th = atan((1.0 * (nk0 / 2 - m)) / (nk0 / 2 - l - 1));
            const int it = (int)floor(th / Dtheta);
            if (it >= 0 && it < N_theta) o->chijj3_k[it + (size_t)n * N_theta] += creal(contrib[q]);
        }
        o->chijj3[n] = creal(s);
    }
    /* Lambda_jj1 + Lambda_jj2 + Lambda_jj3 on the sparse set (:6460-6467) */
    o->chijj_all[0] = o->chijj1[0] + o->chijj2[0];
    for (int j = 1; j < Ns; j++) o->chijj_all[j] = o->chijj1[o->al_freq[j]] + o->chijj2[o->al_freq[j]] + o->chijj3[j];

    free(HG);
    free(HR);
    free(HK);
    free(HK2);
    free(dHn);
    free(dHr);
    free(dK);
    free(dK2);
    free(contrib);
    free(h);
    free(dhG);
    free(Grt);
    free(dGrt);
    free(dek);
    return 0;
}

/* ---------------------------------------------------------------- local G (green.cpp:209-266) */
int orc_gwn_local(orc *o, double complex *out)
{
    if (!o->Self || !o->mu_ok) return -1;
    const int nk = o->nbk, H = o->nbw / 2;
    const long nk8 = (long)nk * (nk + 1) / 2;
    const double nk02 = 4.0 * (nk - 1) * (nk - 1);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < H; j++) {
        const double complex z = I * ((2 * j + 1) * PI * o->p.T);
        double complex g = 0;
        for (int l = 0; l < nk; l++)
            for (int m = 0; m <= l; m++) {
                int w = 8;
                if (l == 0 || l == nk - 1) w /= 2;
                if (m == 0 || m == nk - 1) w /= 2;
                if (m == l) w /= 2;
                const long k = m + (l * (l + 1)) / 2;
                const double complex s = o->Sigma_inf / z + o->Sigma_inf2[k] / (z * z) + o->Sigma_inf3[k] / (z * z * z) +
                                         o->Self[k + (size_t)j * nk8] / (z * z * z * z);
                g += ((double)w / nk02) / (z - o->ek[m + l * nk] + o->mu - s);
            }
        out[j] = g;
    }
    return 0;
}

/* ---------------------------------------------------------------- loss of the gradient tests */
/* mu0 becomes an input: every later set_grid keeps it instead of solving the density */
void orc_fix_mu0(orc *o, double mu0)
{
    o->mu0_fixed = 1;
    o->mu0 = mu0;
}

/* L = sum over the packed wedge of w_k Re Sigma(k, i w_0), Sigma(k,iw_n) = Sigma_inf/z + C2/z^2
   + C3/z^3 + R/z^4 (green.cpp:3039), after the chain chi0 -> Usp, Uch -> V -> Sigma
   (chi.cpp:210-417, 37-206; green.cpp:1673-2031).  Returns NAN if a root is not found. */
double orc_sigma_loss(orc *o, const double *w)
{
    if (!orc_find_Usp_Uch(o)) return NAN;
    orc_vertex(o);
    orc_self_energy(o);
    const long nk8 = (long)o->nbk * (o->nbk + 1) / 2;
    const double complex z = I * (PI * o->p.T);
    double L = 0.0;
    for (long k = 0; k < nk8; k++) {
        const double complex s = o->Sigma_inf / z + o->Sigma_inf2[k] / (z * z) + o->Sigma_inf3[k] / (z * z * z) +
                                 o->Self[k] / (z * z * z * z);
        L += w[k] * creal(s);
    }
    return L;
}

/* The loss terms of tpsc_loss_forward after chi0 -> Usp, Uch -> V -> Sigma (and, for the
   Lambda_jj term, traceSelfG -> mu -> Lambda_jj): bit 1 sum_k w_k Re Sigma(k, i pi T), bit 2
   the same for Im, bit 4 chi_sp(q = (pi, pi), iq_0) = chi0/(1 - Usp chi0/2), bit 8
   sum_{n <= 8} [chijj1 + chijj2](iq_n).  NaN if a root fails. */
double orc_loss(orc *o, const double *w, int terms)
{
    if (!orc_find_Usp_Uch(o)) return NAN;
    orc_vertex(o);
    orc_self_energy(o);
    const long nk8 = (long)o->nbk * (o->nbk + 1) / 2;
    const double complex z = I * (PI * o->p.T);
    double L = 0.0;
    if (terms & 3)
        for (long k = 0; k < nk8; k++) {
            const double complex s = o->Sigma_inf / z + o->Sigma_inf2[k] / (z * z) + o->Sigma_inf3[k] / (z * z * z) +
                                     o->Self[k] / (z * z * z * z);
            if (terms & 1) L += w[k] * creal(s);
            if (terms & 2) L += w[k] * cimag(s);
        }
    if (terms & 4) {
        const double c = o->chiqw[nk8 - 1];
        L += c / (1.0 - o->Usp * c / 2.0);
    }
    if (terms & 8) {
        orc_trace_self_G(o);
        if (!orc_find_mu(o)) return NAN;
        orc_chijj(o, 1);
        for (int n = 0; n <= 8 && n <= o->nwmax; n++) L += o->chijj1[n] + o->chijj2[n];
    }
    return L;
}

/* ---------------------------------------------------------------- spin correlation length */
typedef struct {
    const orc *o;
    double kx0;
} fs_ctx;
static double fs_diag(void *ctx, double kx)
{
    const fs_ctx *f = (const fs_ctx *)ctx;
    return orc_dispk(f->o, kx, kx - f->kx0) - f->o->mu0; /* green0::ekmu_diag, green0.cpp:621-624 */
}

/* green0::find_FS(10001) (green0.cpp:626-675) + chi::vF_hot_spots (chi.cpp:419-447) */
double orc_hot_spot_vF(orc *o)
{
    const int nk = 10001;
    double *FS = dalloc(2 * (size_t)nk);
    for (int j = 0; j < 2 * nk; j++) FS[j] = -1;
    fs_ctx fc = {o, 0.0};
    for (int j = 0; j < nk; j++) {
        fc.kx0 = j * PI / (nk - 1);
        const double init[3] = {0.5 * (fc.kx0 + PI), fc.kx0, PI}, lims[2] = {fc.kx0, PI};
        double root = 0.0;
        if (!orc_find_zero(fs_diag, &fc, init, &root, lims, NULL)) break;
        FS[2 * j] = root;
        FS[2 * j + 1] = root - fc.kx0;
    }
    double kx1 = FS[0], ky1 = FS[1], kx2 = FS[2], ky2 = FS[3];
    int j = 1;
    while ((kx1 + ky1 - PI) * (kx2 + ky2 - PI) > 0 && j < nk - 1) {
        j++;
        kx1 = kx2;
        ky1 = ky2;
        kx2 = FS[2 * j];
        ky2 = FS[2 * j + 1];
    }
    free(FS);
    const double kx = (kx1 + kx2) / 2, ky = (ky1 + ky2) / 2;
    const double gx = orc_grad_ek_x(o, kx, ky), gy = orc_grad_ek_x(o, ky, kx); /* grad_y by symmetry of eps */
    return sqrt(gx * gx + gy * gy);
}

/* chi::sp_corr_length (chi.cpp:451-640) on chi0(q, iq_0) given as the packed wedge c0:
   out = {xi_y, xi_x, xi_as_y, xi_as_x}.  A value the reference reads before setting it
   (chi_sp(qmax_x, 0) already above half maximum) is taken as 0. */
void orc_sp_corr_from(const double *c0, int nbq, int qx, int qy, double chi0max, double Usp, double out[4])
{
#define W(l, m) c0[(m) + ((l) * ((l) + 1)) / 2]
#define SP(x) ((x) / (1.0 - Usp * (x) / 2.0))
    const double chisp_max = chi0max / (1.0 - Usp * chi0max / 2.0), dq = PI / (nbq - 1);
    int l = qx, m = 0;
    double s = SP(W(l, 0)), s2 = 0.0;
    while (s < chisp_max / 2.0 && m <= qy) {
        m++;
        s2 = s;
        s = SP(W(l, m));
    }
    double qa = (m - 1) * dq, qm = qy * dq, qHM = dq * (chisp_max / 2.0 - s2) / (s - s2) + qa;
    out[0] = 1.0 / (qm - qHM);
    m = qy;
    l = m;
    s = SP(W(l, m));
    if (s < chisp_max / 2.0) {
        while (s < chisp_max / 2.0 && l <= qx) {
            l++;
            s2 = s;
            s = SP(W(l, m));
        }
        qa = (l - 1) * dq;
        qm = qx * dq;
        qHM = dq * (chisp_max / 2.0 - s2) / (s - s2) + qa;
        out[1] = 1.0 / (qm - qHM);
    } else
        out[1] = out[0];
    const double r = sqrt(Usp / (2.0 / chi0max - Usp));
    if (qx == nbq - 1 && qy == nbq - 1) {
        const double d2 = 2.0 * (W(nbq - 1, nbq - 2) - chi0max) * (nbq - 1) * (nbq - 1) / (PI * PI);
        out[3] = sqrt(-0.5 * d2 / chi0max) * r;
        out[2] = out[3];
    } else {
        int x = qx, y = qy - 1;
        if (y < 0) y = qy + 1;
        const double cl = W(x, y), cm = W(qx, qy);
        y = qy + 1;
        if (y > x) {
            x = qy + 1;
            y = qx;
        }
        const double cr = W(x, y);
        out[2] = sqrt(-0.5 * (cr - 2 * cm + cl) * (nbq - 1) * (nbq - 1) / (PI * PI) / chi0max) * r;
        x = qx - 1;
        y = qy;
        if (y > x) {
            x = qy;
            y = qx - 1;
            if (y < 0) y = qx + 1;
        }
        const double cl2 = W(x, y);
        x = qx + 1;
        if (x > nbq - 1) x = qx - 1;
        const double cr2 = W(x, qy);
        out[3] = sqrt(-0.5 * (cr2 - 2 * cm + cl2) * (nbq - 1) * (nbq - 1) / (PI * PI) / chi0max) * r;
    }
#undef W
#undef SP
}

/* out = {vF, xi_y, xi_x, xi_as_y, xi_as_x} after orc_find_Usp_Uch (full grid) */
int orc_sp_corr_length(orc *o, double out[5])
{
    if (!o->chiqw || o->Usp == 0.0) return -1;
    out[0] = orc_hot_spot_vF(o);
    orc_sp_corr_from(o->chiqw, o->nbk, o->qmax[0], o->qmax[1], o->chi0max, o->Usp, out + 1);
    return 0;
}
