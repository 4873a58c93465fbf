// tpsc_pade.cpp -- host FP64 analytic continuation: Thiele continued-fraction Pade
// approximants (generique.h:112, generique.cpp:1605-1663) and the optical conductivity
// sigma(w) = Im Lambda_jj(w + i eta) / w from the Matsubara values Lambda_jj(iq_n), continued
// like the bosonic chi_Re_w of chi.cpp:1055-1093.  Plain C++ on the host: no device work.
#include <cmath>
#include <vector>

#include "../../include/tpsc_capi.h"

namespace {

// complex double with the plain (c^2+d^2) division, as the reference's FP64 code computes it
struct cx {
    double re, im;
};
inline cx add(cx a, cx b) { return cx{a.re + b.re, a.im + b.im}; }
inline cx sub(cx a, cx b) { return cx{a.re - b.re, a.im - b.im}; }
inline cx mul(cx a, cx b) { return cx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline cx dvd(cx a, cx b)
{
    const double den = b.re * b.re + b.im * b.im;
    return cx{(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
}
inline cx at(const double* v, int i) { return cx{v[2 * i], v[2 * i + 1]}; }

// coef[j] = g_j(z_j) with g_0 = f and g_m(z) = (coef[m-1] - g_{m-1}(z)) / ((z - z_{m-1}) g_{m-1}(z))
// (pade_recursion, generique.cpp:1650-1663, unrolled for one point at a time)
void pade_coef(const double* z, const double* f, int n, std::vector<cx>& coef)
{
    coef.assign(n, cx{0.0, 0.0});
    for (int j = 0; j < n; j++) {
        cx g = at(f, j);
        const cx zj = at(z, j);
        for (int m = 1; m <= j; m++) g = dvd(sub(coef[m - 1], g), mul(sub(zj, at(z, m - 1)), g));
        coef[j] = g;
    }
}

// forward A/B recurrence of generique::pade (generique.cpp:1630-1648)
cx pade_eval(cx x, int n, const double* z0, const std::vector<cx>& coef)
{
    const cx one{1.0, 0.0};
    cx A = one;
    cx B = (n > 1) ? add(one, mul(sub(x, at(z0, 0)), coef[1])) : one;
    cx q = dvd(A, B);
    for (int j = 3; j <= n; j++) {
        const cx t = mul(sub(x, at(z0, j - 2)), coef[j - 1]);
        A = add(one, dvd(t, A));
        B = add(one, dvd(t, B));
        q = dvd(mul(q, A), B);
    }
    return mul(q, coef[0]);
}

}  // namespace

extern "C" {

int tpsc_pade_coef(const double* z, const double* f, int n, double* coef)
{
    if (!z || !f || !coef || n < 1) return TPSC_E_INVALID;
    std::vector<cx> c;
    pade_coef(z, f, n, c);
    for (int j = 0; j < n; j++) {
        coef[2 * j] = c[j].re;
        coef[2 * j + 1] = c[j].im;
    }
    return TPSC_OK;
}

int tpsc_pade_eval(const double* z, const double* coef, int n, const double* x, int nx, double* out)
{
    if (!z || !coef || !x || !out || n < 1 || nx < 0) return TPSC_E_INVALID;
    std::vector<cx> c(n);
    for (int j = 0; j < n; j++) c[j] = at(coef, j);
    for (int k = 0; k < nx; k++) {
        const cx v = pade_eval(at(x, k), n, z, c);
        out[2 * k] = v.re;
        out[2 * k + 1] = v.im;
    }
    return TPSC_OK;
}

// sigma(w) from Lambda(iq_n) = lam[n], q_n = 2 pi n T, through npade points (see tpsc_capi.h)
int tpsc_pade_conductivity(const double* lam, int nlam, double T, int npade, const double* omega, int nw,
                           double eta, double* sigma)
{
    if (!lam || !omega || !sigma || npade < 2 || npade > nlam || nw < 0 || !(T > 0.0) || !(eta >= 0.0))
        return TPSC_E_INVALID;
    std::vector<double> z(2 * npade), f(2 * npade);
    for (int n = 0; n < npade; n++) {
        z[2 * n] = 0.0;
        z[2 * n + 1] = 2.0 * n * M_PI * T;
        f[2 * n] = lam[n];
        f[2 * n + 1] = 0.0;
    }
    std::vector<cx> c;
    pade_coef(z.data(), f.data(), npade, c);
    for (int k = 0; k < nw; k++) {
        const double w = (omega[k] == 0.0) ? 1.0e-6 : omega[k];
        const cx v = pade_eval(cx{w, eta}, npade, z.data(), c);
        sigma[k] = v.im / w;
    }
    return TPSC_OK;
}

}  // extern "C"
