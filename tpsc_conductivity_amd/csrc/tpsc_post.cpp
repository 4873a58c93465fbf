// tpsc_post.cpp -- host post-processing of a solved parameter point, through the public C ABI
// only (tpsc_get, tpsc_get_scalar, tpsc_get_params, tpsc_find_zero, tpsc_pade_*): the
// infinite-lattice mu0, the Fermi surfaces, the real-frequency self-energy and spectral weight by
// Pade continuation (MDC, EDC, calc_Self_Re_w), and the reference's binary restart files.
// All of it is FP64 host arithmetic over arrays the solve leaves in HBM, as in the reference,
// which runs these steps on the host after find_mu (tpsc_main.cpp:77-101).
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <utility>
#include <vector>

#include "../../include/tpsc_capi.h"

namespace {

using cplx = std::complex<double>;
constexpr double kPi = 3.14159265358979323846;

double hop(const tpsc_params& p) { return p.t != 0.0 ? p.t : 1.0; }

// eps(k) of hamiltonien::dispk (hamiltonien.cpp:119-122), with the hopping t
double dispk(const tpsc_params& p, double kx, double ky)
{
    const double cx = cos(kx), cy = cos(ky), c2x = cos(2 * kx), c2y = cos(2 * ky);
    return -2.0 * hop(p) * (cx + cy) - 4.0 * p.tp * cx * cy - 2.0 * p.tpp * (c2x + c2y) - 4 * p.t3 * (c2x * cy + c2y * cx);
}

long widx(long x, long y) { return y + (x * (x + 1)) / 2; }

double scalar(tpsc_ctx* c, int which)
{
    double v = 0.0;
    tpsc_get_scalar(c, which, &v);
    return v;
}

int field(tpsc_ctx* c, int f, std::vector<double>& v)
{
    const size_t n = tpsc_field_size(c, f) / sizeof(double);
    v.assign(n, 0.0);
    if (!n) return TPSC_E_ORDER;
    return tpsc_get(c, f, v.data(), n * sizeof(double));
}

// Thiele continued fraction through (z[i], f[i]), i < n (generique.cpp:1630-1663), as the
// library's tpsc_pade_coef / tpsc_pade_eval compute it
struct Pade {
    std::vector<double> z, coef;
    int n = 0;
    int fit(const std::vector<cplx>& zz, const std::vector<cplx>& ff, int np)
    {
        n = np;
        z.assign(2 * np, 0.0);
        coef.assign(2 * np, 0.0);
        std::vector<double> f(2 * np);
        for (int i = 0; i < np; i++) {
            z[2 * i] = zz[i].real();
            z[2 * i + 1] = zz[i].imag();
            f[2 * i] = ff[i].real();
            f[2 * i + 1] = ff[i].imag();
        }
        return tpsc_pade_coef(z.data(), f.data(), np, coef.data());
    }
    cplx operator()(cplx x) const
    {
        double xi[2] = {x.real(), x.imag()}, o[2] = {0.0, 0.0};
        tpsc_pade_eval(z.data(), coef.data(), n, xi, 1, o);
        return cplx(o[0], o[1]);
    }
};

// The state of a solved point that the post-processing reads: eps(k) on the quadrant, mu, T and
// Sigma(k, iw_n) for the first nfreq stored frequencies, n = 0, 1, ... ([n*nk8 + k]); with the
// remainder R(k, n) and the moments when asked (FP64 path from the stored R; mixed path from the
// stored assembled Sigma, R = (Sigma - Sinf/z - C2/z^2 - C3/z^3) z^4).
struct Solved {
    tpsc_params p{};
    int nk = 0, nbw = 0;
    long nk8 = 0;
    double T = 0.0, mu = 0.0, mu0 = 0.0, Sinf = 0.0;
    std::vector<double> ek, C2, C3;
    std::vector<cplx> sig, R;
    cplx S(int j, long k) const { return sig[(size_t)j * nk8 + k]; }
};

int load(tpsc_ctx* c, Solved& s, bool need_sigma, int nfreq = 0, bool need_R = false)
{
    if (!c) return TPSC_E_INVALID;
    tpsc_get_params(c, &s.p);
    s.nk = (int)scalar(c, TPSC_S_NK);
    s.nk8 = (long)scalar(c, TPSC_S_NK8);
    s.nbw = 2 * ((int)scalar(c, TPSC_S_NWB) - 1);
    s.T = s.p.T;
    s.mu0 = scalar(c, TPSC_S_MU0);
    if (int rc = field(c, TPSC_F_EK, s.ek)) return rc;
    if (!need_sigma) return TPSC_OK;
    const int st = (int)scalar(c, TPSC_S_STAGES);
    if (!(st & TPSC_STAGE_SIGMA) || !(st & TPSC_STAGE_MU)) return TPSC_E_ORDER;
    s.mu = scalar(c, TPSC_S_MU);
    s.Sinf = scalar(c, TPSC_S_SIGMA_INF);
    std::vector<double> a;
    if (int rc = field(c, TPSC_F_SIGMA_KW, a)) return rc;
    const int H = s.nbw / 2;
    if (nfreq > H) return TPSC_E_INVALID;
    s.sig.resize((size_t)nfreq * s.nk8);
    for (size_t i = 0; i < s.sig.size(); i++) s.sig[i] = cplx(a[2 * i], a[2 * i + 1]);
    if (!need_R) return TPSC_OK;
    if (int rc = field(c, TPSC_F_SIGMA_INF2, s.C2)) return rc;
    if (int rc = field(c, TPSC_F_SIGMA_INF3, s.C3)) return rc;
    s.R.resize((size_t)nfreq * s.nk8);
    if (s.p.precision == TPSC_PREC_FP64) {
        if (int rc = field(c, TPSC_F_SELF, a)) return rc;
        for (size_t i = 0; i < s.R.size(); i++) s.R[i] = cplx(a[2 * i], a[2 * i + 1]);
    } else {
        for (int j = 0; j < nfreq; j++) {
            const cplx z(0.0, (2 * j + 1) * kPi * s.T);
            for (long k = 0; k < s.nk8; k++)
                s.R[(size_t)j * s.nk8 + k] =
                    (s.S(j, k) - s.Sinf / z - s.C2[k] / (z * z) - s.C3[k] / (z * z * z)) * (z * z * z * z);
        }
    }
    return TPSC_OK;
}

// the reference's Sigma(k, z0) on the fit points: stored frequency j at the Matsubara point z0[j]
// (green.cpp:433, 869, 990, 1128 index the stored frequencies by the fit point, not by n[j])
std::vector<cplx> fit_values(const Solved& s, long k, int n)
{
    std::vector<cplx> f(n);
    for (int j = 0; j < n; j++) f[j] = s.S(j, k);
    return f;
}

double akw_of(cplx z, double ekmu, cplx sw) { return -std::imag(1.0 / (z - ekmu - sw)) / kPi; }

// A(k, z) along the diagonals x = l + m, y = m (m < nk - l): the first m of largest A > 0
// (green0.cpp:77-100, green.cpp:2262-2325)
template <typename Fn>
void argmax_diagonals(int nk, const Fn& akw, int* midx, double* amax)
{
    for (int l = 0; l < nk; l++) {
        double best = 0.0;
        int bm = 0;
        for (int m = 0; m < nk - l; m++) {
            const double a = akw(l, m);
            if (a > best) {
                best = a;
                bm = m;
            }
        }
        if (midx) midx[l] = bm;
        if (amax) amax[l] = best;
    }
}

// -------------------------------------------------- infinite-lattice density (green0.cpp:221-236)
// (2/pi^2) int_0^pi int_0^pi f(eps(k) - mu) dkx dky by nested adaptive Gauss-Kronrod (7, 15)
// quadrature: a panel is accepted when |K15 - G7| <= tol, else halved with tol/2 on each half.
// The 15-point rule is exact to degree 22, so the accepted panels are accurate far below tol.
const double kXK[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                       0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                       0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                       0.207784955007898467600689403773245, 0.000000000000000000000000000000000};
const double kWK[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                       0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                       0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                       0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
const double kWG[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                       0.381830050505118944950369775488975, 0.417959183673469387755102040816327};

template <typename Fn>
double gk15(const Fn& f, double a, double b, double* err)
{
    const double c = 0.5 * (a + b), h = 0.5 * (b - a);
    const double fc = f(c);
    double k = fc * kWK[7], g = fc * kWG[3];
    for (int i = 0; i < 7; i++) {
        const double x = h * kXK[i];
        const double s = f(c - x) + f(c + x);
        k += kWK[i] * s;
        if (i % 2 == 1) g += kWG[i / 2] * s;
    }
    *err = fabs((k - g) * h);
    return k * h;
}

template <typename Fn>
double adapt(const Fn& f, double a, double b, double tol, int depth)
{
    double err = 0.0;
    const double v = gk15(f, a, b, &err);
    if (err <= tol || depth >= 40) return v;
    const double c = 0.5 * (a + b);
    return adapt(f, a, c, 0.5 * tol, depth + 1) + adapt(f, c, b, 0.5 * tol, depth + 1);
}

double fermi(double x, double T)
{
    if (x > 0) {
        const double e = exp(-x / T);
        return e / (1.0 + e);
    }
    return 1.0 / (exp(x / T) + 1.0);
}

double density_infinite(const tpsc_params& p, double mu)
{
    const double tol = 1e-13;
    auto inner = [&](double ky) {
        return adapt([&](double kx) { return fermi(dispk(p, kx, ky) - mu, p.T); }, 0.0, kPi, tol, 0);
    };
    return 2.0 / (kPi * kPi) * adapt(inner, 0.0, kPi, tol, 0);
}

struct RootCtx {
    const tpsc_params* p;
    double a;      // the target density, or kx0 of a Fermi-surface trace
    double level;  // mu0 of the Fermi-surface traces
};

double dens_fn(double mu, void* u)
{
    const RootCtx* r = (const RootCtx*)u;
    return density_infinite(*r->p, mu) - r->a;
}

double fs_fn(double kx, void* u)
{
    const RootCtx* r = (const RootCtx*)u;
    return dispk(*r->p, kx, kx - r->a) - r->level;
}

// --------------------------------------------------------------------------- file helpers
std::string path(const char* dir, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
std::string path(const char* dir, const char* fmt, ...)
{
    char name[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(name, sizeof name, fmt, ap);
    va_end(ap);
    const char* base = (strncmp(name, "./", 2) == 0) ? name + 2 : name;
    return (dir && *dir) ? std::string(dir) + "/" + base : std::string(base);
}

}  // namespace

extern "C" {

int tpsc_mu0_infinite(tpsc_ctx* c, double* mu0)
{
    if (!c || !mu0) return TPSC_E_INVALID;
    tpsc_params p;
    if (int rc = tpsc_get_params(c, &p)) return rc;
    if (p.n == 1.0 && p.tp == 0 && p.tpp == 0) {  // green0.cpp:268-272
        *mu0 = 0.0;
        return TPSC_OK;
    }
    RootCtx r{&p, p.n, 0.0};
    const double t8 = 8.0 * fabs(hop(p));
    const double init[3] = {0.0, -t8, t8};  // green0.cpp:277
    double root = 0.0;
    const int rc = tpsc_find_zero(dens_fn, &r, init, nullptr, &root, nullptr);
    if (rc) return rc;
    *mu0 = root;
    return TPSC_OK;
}

int tpsc_fermi_surface(tpsc_ctx* c, int nk, double* kxy, int* count)
{
    if (!c || !kxy || nk < 2) return TPSC_E_INVALID;
    tpsc_params p;
    if (int rc = tpsc_get_params(c, &p)) return rc;
    const double mu0 = scalar(c, TPSC_S_MU0);
    for (int j = 0; j < 2 * nk; j++) kxy[j] = -1.0;
    int found = 0;
    for (int j = 0; j < nk; j++) {
        const double kx0 = j * kPi / (nk - 1);
        RootCtx r{&p, kx0, mu0};
        const double init[3] = {0.5 * (kx0 + kPi), kx0, kPi}, lims[2] = {kx0, kPi};
        double root = 0.0;
        if (tpsc_find_zero(fs_fn, &r, init, lims, &root, nullptr)) break;
        kxy[2 * j] = root;
        kxy[2 * j + 1] = root - kx0;
        found++;
    }
    if (count) *count = found;
    return TPSC_OK;
}

int tpsc_fs_lattice(tpsc_ctx* c, int kind, int* m_idx, double* Ak)
{
    if (!c || kind < 0 || kind > 2) return TPSC_E_INVALID;
    Solved s;
    if (int rc = load(c, s, kind != 0, kind ? 4 : 0)) return rc;
    const int nk = s.nk;
    if (kind == 0) {  // green0::find_FS_PBC, eta = 1e-5 (green0.cpp:59)
        const cplx z(0.0, 1.0e-5);
        argmax_diagonals(nk, [&](int l, int m) { return akw_of(z, s.ek[(l + m) + m * nk] - s.mu0, 0.0); }, m_idx, Ak);
        return TPSC_OK;
    }
    // green::find_FS_inter: 3-point polynomial (kind 1) or 4-point Pade (kind 2) of Sigma(k, iw_j),
    // j < 4, at z = i 1e-6 (green.cpp:2228-2313)
    const cplx z(0.0, 1.0e-6);
    cplx z0[4];
    for (int j = 0; j < 4; j++) z0[j] = cplx(0.0, (2 * j + 1) * kPi * s.T);
    auto akw = [&](int l, int m) {
        const long x = l + m, y = m, k = widx(x, y);
        const double ekmu = s.ek[x + y * nk] - s.mu;
        cplx sw;
        if (kind == 2) {
            Pade P;
            P.fit(std::vector<cplx>(z0, z0 + 4), fit_values(s, k, 4), 4);
            sw = P(z);
        } else {
            const cplx a = s.S(0, k), b = s.S(1, k), d = s.S(2, k);
            sw = a * ((z - z0[1]) / (z0[0] - z0[1])) * ((z - z0[2]) / (z0[0] - z0[2])) +
                 b * ((z - z0[0]) / (z0[1] - z0[0])) * ((z - z0[2]) / (z0[1] - z0[2])) +
                 d * ((z - z0[0]) / (z0[2] - z0[0])) * ((z - z0[1]) / (z0[2] - z0[1]));
        }
        return akw_of(z, ekmu, sw);
    };
    argmax_diagonals(nk, akw, m_idx, Ak);
    return TPSC_OK;
}

int tpsc_mdc(tpsc_ctx* c, const int* n, int Nn, double w, const double* eta, int Neta, int use_pade, double* self_w,
             double* akw)
{
    if (!c || !n || !eta || Neta < 1) return TPSC_E_INVALID;
    if (Nn < 2) Nn = 2;  // green.cpp:822
    if (!use_pade && Nn < 3) return TPSC_E_INVALID;
    Solved s;
    if (int rc = load(c, s, true, Nn)) return rc;
    std::vector<cplx> z0(Nn);
    for (int j = 0; j < Nn; j++) z0[j] = cplx(0.0, (2 * n[j] + 1) * kPi * s.T);
    const int nk = s.nk;
    for (int l = 0; l < nk; l++)
        for (int m = 0; m <= l; m++) {
            const long k = widx(l, m);
            const double ekmu = s.ek[m + l * nk] - s.mu;
            const std::vector<cplx> f = fit_values(s, k, Nn);
            Pade P;
            if (use_pade) P.fit(z0, f, Nn);
            for (int e = 0; e < Neta; e++) {
                const cplx z(w, eta[e]);
                cplx sw;
                if (use_pade)
                    sw = P(z);
                else {  // Lagrange polynomial through 4 (Nn >= 4) or 3 points (green.cpp:870-882)
                    const int np = Nn >= 4 ? 4 : 3;
                    sw = 0.0;
                    for (int i = 0; i < np; i++) {
                        cplx t = f[i];
                        for (int j = 0; j < np; j++)
                            if (j != i) t *= (z - z0[j]) / (z0[i] - z0[j]);
                        sw += t;
                    }
                }
                if (self_w) {
                    self_w[2 * (k * Neta + e)] = sw.real();
                    self_w[2 * (k * Neta + e) + 1] = sw.imag();
                }
                if (akw) akw[k * Neta + e] = akw_of(z, ekmu, sw);
            }
        }
    return TPSC_OK;
}

int tpsc_edc(tpsc_ctx* c, const int* kx, const int* ky, int Nk, const double* w, int nw0, int m_freq, int NP_max,
             const double* eta, int Neta, double* spectr)
{
    if (!c || !kx || !ky || Nk < 1 || !w || nw0 < 1 || !eta || Neta < 1 || !spectr || m_freq < 0 || NP_max < 1)
        return TPSC_E_INVALID;
    if (w[0] > w[nw0 - 1]) return TPSC_E_INVALID;  // green.cpp:1077-1081
    Solved s;
    if (int rc = load(c, s, false)) return rc;
    // green::create_non_uniform_ind_freq (green.cpp:556-583)
    const int N1 = s.nbw / (1 << (m_freq + 1)), N2 = N1 / 2, nf = N1 + m_freq * N2;
    if (N1 < 1 || (m_freq > 0 && N2 < 1)) return TPSC_E_INVALID;
    std::vector<cplx> z0(nf);
    for (int j = 0; j < nf; j++) {
        int ind = j;
        if (j >= N1) {
            const int lj = (j - N1) / N2, p2 = 1 << lj;
            ind = p2 * N1 + p2 * 2 * ((j - N1) % N2);
        }
        z0[j] = cplx(0.0, (2 * ind + 1) * kPi * s.T);
    }
    const int NP = nf < NP_max ? nf : NP_max;
    if (NP > s.nbw / 2) return TPSC_E_INVALID;  // green.cpp:1083-1087
    if (int rc = load(c, s, true, NP)) return rc;
    const int nk = s.nk, Nkk = 2 * (nk - 1);
    for (int p = 0; p < Nk; p++) {
        int k0 = kx[p], k1 = ky[p];  // green.cpp:1095-1099
        while (k0 < 0) k0 += Nkk;
        while (k1 < 0) k1 += Nkk;
        if (k0 > nk - 1) k0 = Nkk - k0;
        if (k1 > nk - 1) k1 = Nkk - k1;
        if (k0 < 0 || k1 < 0) return TPSC_E_INVALID;
        long x = k0, y = k1;
        if (y > x) std::swap(x, y);
        const long k = widx(x, y);
        Pade P;
        P.fit(z0, fit_values(s, k, NP), NP);
        const double ekmu = s.ek[k1 + k0 * nk] - s.mu;  // ek_val (hamiltonien.cpp:107-116)
        double* out = spectr + (size_t)p * nw0 * Neta;
        for (int i = 0; i < nw0; i++)
            for (int e = 0; e < Neta; e++) {
                const cplx sw = P(cplx(w[i], eta[e]));
                out[e + i * Neta] = -std::imag(1.0 / (cplx(w[i] - ekmu, eta[e]) - sw)) / kPi;
            }
    }
    return TPSC_OK;
}

int tpsc_self_real_freq(tpsc_ctx* c, const int* kv, const int* n, int Nn, const double* w, int Nw, const double* eta,
                        int Neta, double* self_w, double* akw, double* self_jfrac, double* akw_jfrac)
{
    if (!c || !kv || !n || Nn < 4 || !w || Nw < 1 || !eta || Neta < 1) return TPSC_E_INVALID;
    Solved s;
    if (int rc = load(c, s, true, Nn, true)) return rc;
    const int nk = s.nk;
    if (kv[0] < 0 || kv[1] < 0 || kv[0] >= nk || kv[1] >= nk) return TPSC_E_INVALID;
    long x = kv[0], y = kv[1];
    if (y > x) std::swap(x, y);
    const long k = widx(x, y);
    std::vector<cplx> z0(Nn);
    for (int j = 0; j < Nn; j++) z0[j] = cplx(0.0, (2 * n[j] + 1) * kPi * s.T);
    // Sigma_inf/z + C2/z^2 + C3/z^3 + R/z^4 on the fit points z = i w (green.cpp:428-433), in the
    // real form -C2/w^2 + Re R/w^4 + i(-Sigma_inf/w + C3/w^3 + Im R/w^4) of the Sigma(k, iw_n) copy-out
    std::vector<cplx> f(Nn);
    for (int j = 0; j < Nn; j++) {
        const double wn = (2 * n[j] + 1) * kPi * s.T, w2 = wn * wn, w3 = w2 * wn, w4 = w2 * w2;
        const cplx r = s.R[(size_t)j * s.nk8 + k];
        f[j] = cplx(-s.C2[k] / w2 + r.real() / w4, -s.Sinf / wn + s.C3[k] / w3 + r.imag() / w4);
    }
    Pade P;
    P.fit(z0, f, Nn);
    const double ekmu = s.ek[kv[1] + kv[0] * nk] - s.mu;
    for (int i = 0; i < Nw; i++)
        for (int e = 0; e < Neta; e++) {
            const cplx z(w[i], eta[e]);
            const cplx sw = P(z);
            if (self_w) {
                self_w[2 * (i * Neta + e)] = sw.real();
                self_w[2 * (i * Neta + e) + 1] = sw.imag();
            }
            if (akw) akw[i * Neta + e] = -std::imag((1.0 / kPi) / (z - ekmu - sw));
        }
    // Sigma = a1 / (z + b1 + S(z)), S the Pade of the transformed remainder on Nn/8 (>= 4) points
    // (green.cpp:494-542)
    const double c1 = s.Sinf, c2 = s.C2[k], c3 = s.C3[k];
    const double a1 = c1, b1 = -c2 / c1;
    int Nn2 = Nn / 8;
    if (Nn2 < 4) Nn2 = 4;
    std::vector<cplx> g(Nn2);
    for (int j = 0; j < Nn2; j++) {
        const cplx z = z0[j], fk = s.R[(size_t)j * s.nk8 + k];
        g[j] = (b1 * fk + (fk + b1 * c3) * z + (c3 - a1 * b1 * b1) * z * z) / (-fk - c3 * z + a1 * b1 * z * z - a1 * z * z * z);
    }
    Pade Q;
    Q.fit(std::vector<cplx>(z0.begin(), z0.begin() + Nn2), g, Nn2);
    for (int i = 0; i < Nw; i++)
        for (int e = 0; e < Neta; e++) {
            const cplx z(w[i], eta[e]);
            const cplx sw = a1 / (z + b1 + Q(z));
            if (self_jfrac) {
                self_jfrac[2 * (i * Neta + e)] = sw.real();
                self_jfrac[2 * (i * Neta + e) + 1] = sw.imag();
            }
            if (akw_jfrac) akw_jfrac[i * Neta + e] = -std::imag((1.0 / kPi) / (z - ekmu - sw));
        }
    return TPSC_OK;
}

// green::save_Self_bin (green.cpp:1496-1514): Self_kw_array as the reference holds it after
// calc_Self_FFT_spline, (Nw/2 + 1) x nk8 complex values: R(k, n) for n < Nw/2, then the row the
// remainder pass leaves untouched, (Sigma(k, tau_{Nw/2}), Sigma(k, tau_Nw)) (green.cpp:1872-1885)
int tpsc_save_self_bin(tpsc_ctx* c, const char* dir)
{
    if (!c) return TPSC_E_INVALID;
    tpsc_params p;
    tpsc_get_params(c, &p);
    if (p.precision != TPSC_PREC_FP64) return TPSC_E_INVALID;  // the mixed path keeps no remainder
    if (!((int)scalar(c, TPSC_S_STAGES) & TPSC_STAGE_SIGMA)) return TPSC_E_ORDER;
    std::vector<double> R, St;
    if (int rc = field(c, TPSC_F_SELF, R)) return rc;
    if (int rc = field(c, TPSC_F_SIGMA_TAU, St)) return rc;
    const long nk8 = (long)scalar(c, TPSC_S_NK8);
    const int nwb = (int)scalar(c, TPSC_S_NWB), nbw = 2 * (nwb - 1), nk0 = 2 * ((int)scalar(c, TPSC_S_NK) - 1);
    const int H = nbw / 2;
    R.resize(2 * (size_t)(H + 1) * nk8);
    for (long k = 0; k < nk8; k++) {
        R[2 * ((size_t)H * nk8 + k)] = St[(size_t)H * nk8 + k];
        R[2 * ((size_t)H * nk8 + k) + 1] = St[(size_t)nbw * nk8 + k];
    }
    std::ofstream f(path(dir, "./self_bin_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat", p.U, p.tp, p.tpp, p.t3, p.n,
                         p.T, nk0, nbw),
                    std::ios::out | std::ios::binary);
    if (!f) return TPSC_E_INVALID;
    f.write((const char*)R.data(), R.size() * sizeof(double));
    return f ? TPSC_OK : TPSC_E_INVALID;
}

// chi0::save_chi (chi0.cpp:61-96): int nbq, int nwmax = Nw/2, then chi0(q, iq_n), n = 0..nwmax,
// each row the nk8 wedge values
int tpsc_save_chi0_bin(tpsc_ctx* c, const char* dir)
{
    if (!c) return TPSC_E_INVALID;
    tpsc_params p;
    tpsc_get_params(c, &p);
    if (!((int)scalar(c, TPSC_S_STAGES) & TPSC_STAGE_CHI0)) return TPSC_E_ORDER;
    std::vector<double> chi;
    if (int rc = field(c, TPSC_F_CHI0, chi)) return rc;
    const int nbq = (int)scalar(c, TPSC_S_NK), nwmax = (int)scalar(c, TPSC_S_NWB) - 1;
    std::ofstream f(path(dir, "./chi0_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%d_%d.dat", p.tp, p.tpp, p.t3, p.n, p.T,
                         2 * (nbq - 1), 2 * nwmax),
                    std::ios::out | std::ios::binary);
    if (!f) return TPSC_E_INVALID;
    f.write((const char*)&nbq, sizeof nbq);
    f.write((const char*)&nwmax, sizeof nwmax);
    f.write((const char*)chi.data(), chi.size() * sizeof(double));
    return f ? TPSC_OK : TPSC_E_INVALID;
}

// green::load_Self_bin (green.cpp:1516-1540): the (Nw/2 + 1) * nk8 complex values of a
// self_bin_* file of an N x N, Nw grid into out (interleaved re/im)
int tpsc_read_self_bin(const char* file, int N, int Nw, double* out)
{
    if (!file || !out || N < 4 || Nw < 2) return TPSC_E_INVALID;
    const long nk = N / 2 + 1, nk8 = nk * (nk + 1) / 2;
    const size_t n = 2 * (size_t)(Nw / 2 + 1) * nk8;
    std::ifstream f(file, std::ios::in | std::ios::binary);
    if (!f) return TPSC_E_INVALID;
    f.read((char*)out, n * sizeof(double));
    return (f && (size_t)f.gcount() == n * sizeof(double)) ? TPSC_OK : TPSC_E_INVALID;
}

// chi0::loadChi0 (chi0.cpp:99-156): the header (nbq, nwmax) and, when cap (doubles) allows,
// the (nwmax + 1) * nbq (nbq + 1)/2 values of chi0(q, iq_n)
int tpsc_read_chi0_bin(const char* file, int* nbq, int* nwmax, double* out, size_t cap)
{
    if (!file || !nbq || !nwmax) return TPSC_E_INVALID;
    std::ifstream f(file, std::ios::in | std::ios::binary);
    if (!f) return TPSC_E_INVALID;
    f.read((char*)nbq, sizeof *nbq);
    f.read((char*)nwmax, sizeof *nwmax);
    if (!f || *nbq < 1 || *nwmax < 0) return TPSC_E_INVALID;
    const size_t n = (size_t)(*nwmax + 1) * ((size_t)*nbq * (*nbq + 1) / 2);
    if (!out) return TPSC_OK;
    if (cap < n) return TPSC_E_INVALID;
    f.read((char*)out, n * sizeof(double));
    return (f && (size_t)f.gcount() == n * sizeof(double)) ? TPSC_OK : TPSC_E_INVALID;
}

}  // extern "C"
