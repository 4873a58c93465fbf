"""numpy restatement of the reference's post-processing of a solved point (test infrastructure
only: tests/ import it as the checker of the library's host post-processing).

Functions and the reference code they follow:
* mu0_infinite        green0::find_mu_inf_syst / ddensity_mu   green0.cpp:221-236,266-287
* fermi_surface       green0::find_FS (continuum, along kx - ky = const)   green0.cpp:626-676
* fs_lattice          green0::find_FS_PBC                      green0.cpp:51-105
* fs_inter            green::find_FS_inter                      green.cpp:2210-2341
* sigma_ikn           Sigma(k, iw_n) from its remainder and moments   green.cpp:3039
* mdc                 green::MDC                                green.cpp:814-1068
* non_uniform_freqs   green::create_non_uniform_ind_freq        green.cpp:556-583
* edc                 green::EDC + calc_spectral_weight         green.cpp:589-700,1073-1137
* self_real_freq      green::calc_Self_Re_w                     green.cpp:408-551
The Pade continued fraction is the oracle's own (oracle.pade, generique.cpp:1630-1663).
Reference conventions kept on purpose: the Sigma values fed to the Pade fits are the stored
frequencies j = 0, 1, ... of Self_kw_array whatever the Matsubara indices n[j] of the fit
points (green.cpp:433, 869, 990, 1128); ek is read as ek[x + y*nbk] along the diagonals of
find_FS_PBC / find_FS_inter (eps is symmetric in kx, ky).
"""
import numpy as np
from scipy import integrate, optimize

import oracle

PI = np.pi


def dispersion(kx, ky, tp, tpp, t3, t=1.0):
    cx, cy, c2x, c2y = np.cos(kx), np.cos(ky), np.cos(2 * kx), np.cos(2 * ky)
    return -2 * t * (cx + cy) - 4 * tp * cx * cy - 2 * tpp * (c2x + c2y) - 4 * t3 * (c2x * cy + c2y * cx)


def fermi(x, T):
    x = np.asarray(x, dtype=float)
    return np.where(x > 0, np.exp(-np.abs(x) / T) / (1 + np.exp(-np.abs(x) / T)), 1 / (1 + np.exp(-np.abs(x) / T)))


def density_infinite(mu, T, tp, tpp, t3, t=1.0, tol=1e-13):
    """(2/pi^2) int_0^pi int_0^pi f(eps(k) - mu) dkx dky by nested QUADPACK."""
    def inner(ky):
        return integrate.quad(lambda kx: float(fermi(dispersion(kx, ky, tp, tpp, t3, t) - mu, T)), 0.0, PI,
                              epsabs=tol, epsrel=tol, limit=400)[0]

    return 2.0 / PI ** 2 * integrate.quad(inner, 0.0, PI, epsabs=tol, epsrel=tol, limit=400)[0]


def mu0_infinite(n, T, tp, tpp, t3, t=1.0):
    """mu0 of the infinite lattice: root of density_infinite(mu) - n in [-8, 8] (green0.cpp:266-287);
    0 at half filling without t', t''."""
    if n == 1.0 and tp == 0 and tpp == 0 and t3 == 0:
        return 0.0
    return optimize.brentq(lambda m: density_infinite(m, T, tp, tpp, t3, t) - n, -8.0 * abs(t), 8.0 * abs(t),
                           xtol=1e-14, rtol=1e-15)


def fermi_surface(mu0, tp, tpp, t3, nk=10001, t=1.0):
    """For kx0 = j pi/(nk-1), a root kx in [kx0, pi] of eps(kx, kx - kx0) = mu0, searched from the
    midpoint inside the limits [kx0, pi] (green0.cpp:643-674): the half [mid, pi] when eps - mu0
    changes sign over it, else [kx0, mid]; the list ends at the first j with neither."""
    out = []
    for j in range(nk):
        kx0 = j * PI / (nk - 1)

        def f(kx):
            return dispersion(kx, kx - kx0, tp, tpp, t3, t) - mu0

        # |f| <= 1e-15 is a root (the tolerance of the reference's search with limits)
        mid = 0.5 * (kx0 + PI)
        a, b = f(kx0), f(PI)
        r = None
        if abs(a) <= 1e-15:
            r = kx0
        elif abs(b) <= 1e-15:
            r = PI
        elif kx0 < mid < PI:
            fm = f(mid)
            if abs(fm) <= 1e-15:
                r = mid
            elif fm * b < 0:
                r = optimize.brentq(f, mid, PI, xtol=1e-15, rtol=1e-15)
            elif fm * a < 0:
                r = optimize.brentq(f, kx0, mid, xtol=1e-15, rtol=1e-15)
        if r is None:
            break
        out.append((r, r - kx0))
    return np.array(out).reshape(-1, 2)


def _argmax_diag(nbk, akw):
    """per diagonal l: the first m (m < nbk - l) with the largest akw(l, m) > 0; (m, A) or (0, 0)"""
    idx = np.zeros(nbk, dtype=int)
    amax = np.zeros(nbk)
    for l in range(nbk):
        for m in range(nbk - l):
            a = akw(l, m)
            if a > amax[l]:
                amax[l] = a
                idx[l] = m
    return idx, amax


def fs_lattice(ek, nbk, mu0, eta=1e-5):
    """find_FS_PBC: A0(k, 0) = -Im 1/(i eta - (eps - mu0))/pi along the diagonals x = l + m, y = m."""
    def akw(l, m):
        ekmu = ek[(l + m) + m * nbk] - mu0
        return -(1.0 / (1j * eta - ekmu)).imag / PI

    return _argmax_diag(nbk, akw)


def sigma_ikn(R, Sinf, C2, C3, T, nfreq):
    """Sigma(k, iw_j), j < nfreq, from the remainder R[j, k] and the moments (green.cpp:3039) in
    complex arithmetic"""
    out = np.empty((nfreq, R.shape[1]), dtype=complex)
    for j in range(nfreq):
        z = 1j * (2 * j + 1) * PI * T
        out[j] = Sinf / z + C2 / (z * z) + C3 / (z * z * z) + R[j] / (z * z * z * z)
    return out


def widx(x, y):
    return y + x * (x + 1) // 2


def fs_inter(ek, nbk, mu, sig, T, eta=1e-6):
    """find_FS_inter: per diagonal, A(k, i eta) from the 3-point polynomial and from the 4-point
    Pade of Sigma(k, iw_j), j < 4; returns ((m_poly, A_poly), (m_pade, A_pade))."""
    z0 = 1j * (2 * np.arange(4) + 1) * PI * T
    z = 1j * eta
    poly, pade = {}, {}
    for l in range(nbk):
        for m in range(nbk - l):
            x, y = l + m, m
            ekmu = ek[x + y * nbk] - mu
            s = sig[:4, widx(x, y)]
            sp = oracle.pade(s, z0, np.array([z]))[0]
            pade[(l, m)] = -(1.0 / (z - ekmu - sp)).imag / PI
            spoly = (s[0] * ((z - z0[1]) / (z0[0] - z0[1])) * ((z - z0[2]) / (z0[0] - z0[2]))
                     + s[1] * ((z - z0[0]) / (z0[1] - z0[0])) * ((z - z0[2]) / (z0[1] - z0[2]))
                     + s[2] * ((z - z0[0]) / (z0[2] - z0[0])) * ((z - z0[1]) / (z0[2] - z0[1])))
            poly[(l, m)] = -(1.0 / (z - ekmu - spoly)).imag / PI
    return _argmax_diag(nbk, lambda l, m: poly[(l, m)]), _argmax_diag(nbk, lambda l, m: pade[(l, m)])


def _lagrange(s, z0, z):
    if len(z0) >= 4:
        terms = 4
    else:
        terms = 3
    out = 0.0
    for i in range(terms):
        p = s[i]
        for j in range(terms):
            if j != i:
                p = p * ((z - z0[j]) / (z0[i] - z0[j]))
        out = out + p
    return out


def mdc(ek, nbk, mu, sig_of, nidx, T, w, etas, use_pade=True):
    """green::MDC at energy w: Sigma(k, w + i eta) and A(k, w + i eta) for every wedge point (l, m <= l),
    in the order of the reference's loops.  sig_of(j, k) = Sigma(k, iw_j) of stored frequency j.
    Returns (self [nk8, neta] complex, akw [nk8, neta])."""
    Nn = max(len(nidx), 2)
    z0 = 1j * (2 * np.asarray(nidx[:Nn]) + 1) * PI * T
    nk8 = nbk * (nbk + 1) // 2
    selfw = np.zeros((nk8, len(etas)), dtype=complex)
    akw = np.zeros((nk8, len(etas)))
    for l in range(nbk):
        for m in range(l + 1):
            k = widx(l, m)
            ekmu = ek[m + l * nbk] - mu
            s = np.array([sig_of(j, k) for j in range(Nn)])
            for e, eta in enumerate(etas):
                z = w + 1j * eta
                sw = oracle.pade(s, z0, np.array([z]))[0] if use_pade else _lagrange(s, z0, z)
                selfw[k, e] = sw
                akw[k, e] = -(1.0 / (z - ekmu - sw)).imag / PI
    return selfw, akw


def non_uniform_freqs(nbw, m):
    """green::create_non_uniform_ind_freq"""
    N1 = nbw // 2 ** (m + 1)
    N2 = N1 // 2
    nf = N1 + m * N2
    out = []
    for j in range(nf):
        if j < N1:
            out.append(j)
        else:
            lj = (j - N1) // N2
            p = 2 ** lj
            out.append(p * N1 + p * 2 * ((j - N1) % N2))
    return np.array(out, dtype=int)


def ek_val(ek, nbk, kx, ky):
    Nk = 2 * (nbk - 1)
    while kx < 0:
        kx += Nk
    while ky < 0:
        ky += Nk
    if kx > nbk - 1:
        kx = Nk - kx
    if ky > nbk - 1:
        ky = Nk - ky
    return ek[ky + kx * nbk]


def edc(ek, nbk, mu, sig_of, T, nbw, kxs, kys, w, m_freq, np_max, etas):
    """green::EDC(kx, ky, Nk, w, nw0, m_freq, NP_max, eta, N_eta): spectral weight [point, w, eta]"""
    ind = non_uniform_freqs(nbw, m_freq)
    NP = min(len(ind), np_max)
    z0 = 1j * (2 * ind[:NP] + 1) * PI * T
    Nk = 2 * (nbk - 1)
    out = np.zeros((len(kxs), len(w), len(etas)))
    for p, (kx, ky) in enumerate(zip(kxs, kys)):
        k = [kx, ky]
        while k[0] < 0:
            k[0] += Nk
        while k[1] < 0:
            k[1] += Nk
        if k[0] > nbk - 1:
            k[0] = Nk - k[0]
        if k[1] > nbk - 1:
            k[1] = Nk - k[1]
        x, y = k
        if y > x:
            x, y = y, x
        s = np.array([sig_of(n, widx(x, y)) for n in range(NP)])
        ekmu = ek_val(ek, nbk, k[0], k[1]) - mu
        for i, wi in enumerate(w):
            for e, eta in enumerate(etas):
                sw = oracle.pade(s, z0, np.array([wi + 1j * eta]))[0]
                out[p, i, e] = -(1.0 / ((wi - ekmu) + 1j * eta - sw)).imag / PI
    return out


def self_real_freq(ek, nbk, mu, Sinf, C2, C3, R_of, T, k, nidx, w, etas):
    """green::calc_Self_Re_w(k, n, Nn, w, Nw, eta, Neta): Sigma(w + i eta) and A from the Nn-point
    Pade, then from the continued-fraction form Sigma = a1/(z + b1 + S(z)) with S the Pade of the
    transformed data on Nn/8 (at least 4) points.  Returns (self, akw, self_jfrac, akw_jfrac)."""
    x, y = k
    if y > x:
        x, y = y, x
    kk = widx(x, y)
    Nn = len(nidx)
    z0 = 1j * (2 * np.asarray(nidx) + 1) * PI * T
    # Sinf/z + C2/z^2 + C3/z^3 + R/z^4 at z = i w in real arithmetic (green.cpp:428-433)
    func = []
    for j in range(Nn):
        wn = (2 * nidx[j] + 1) * PI * T
        w2 = wn * wn
        w3, w4 = w2 * wn, w2 * w2
        r = R_of(j, kk)
        func.append(complex(-C2[kk] / w2 + r.real / w4, -Sinf / wn + C3[kk] / w3 + r.imag / w4))
    func = np.array(func)
    ekmu = ek_val(ek, nbk, k[0], k[1]) - mu
    sw = np.zeros((len(w), len(etas)), dtype=complex)
    ak = np.zeros((len(w), len(etas)))
    for i, wi in enumerate(w):
        for e, eta in enumerate(etas):
            z = wi + 1j * eta
            s = oracle.pade(func, z0, np.array([z]))[0]
            sw[i, e] = s
            ak[i, e] = -((1.0 / PI) / (z - ekmu - s)).imag
    c1, c2, c3 = Sinf, C2[kk], C3[kk]
    a1, b1 = c1, -c2 / c1
    Nn2 = max(Nn // 8, 4)
    f2 = []
    for j in range(Nn2):
        z = z0[j]
        fk = R_of(j, kk)
        f2.append((b1 * fk + (fk + b1 * c3) * z + (c3 - a1 * b1 * b1) * z * z) / (-fk - c3 * z + a1 * b1 * z * z - a1 * z * z * z))
    f2 = np.array(f2)
    sj = np.zeros_like(sw)
    aj = np.zeros_like(ak)
    for i, wi in enumerate(w):
        for e, eta in enumerate(etas):
            z = wi + 1j * eta
            sub = oracle.pade(f2, z0[:Nn2], np.array([z]))[0]
            s = a1 / (z + b1 + sub)
            sj[i, e] = s
            aj[i, e] = -((1.0 / PI) / (z - ekmu - s)).imag
    return sw, ak, sj, aj
