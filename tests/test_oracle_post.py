"""Pins the oracle's restatement of the reference's post-processing (oracle/post.py) by what the
mathematics fixes: particle-hole symmetry and the finite-lattice limit of the infinite-lattice
density, the closed-form Fermi surface at half filling, and exact recovery of rational
self-energies by the Pade and polynomial continuations (MDC, EDC, calc_Self_Re_w, find_FS_inter)."""
import numpy as np
import pytest

import oracle
import post

PI = np.pi


def test_infinite_density_is_one_at_mu_zero_by_particle_hole_symmetry():
    for T in (0.05, 0.2):
        assert abs(post.density_infinite(0.0, T, 0.0, 0.0, 0.0) - 1.0) < 1e-12


def test_infinite_mu0_is_the_limit_of_the_finite_lattice_mu0(oracle_mod):
    # at T = 0.25 the lattice sum of the Fermi function converges to the Brillouin-zone integral
    # faster than any power of 1/N: N = 128 is already at the rounding level
    n, T, tp, tpp = 0.85, 0.25, -0.2, 0.1
    o = oracle_mod.Oracle(128, 32, n, T, 2.0, tp=tp, tpp=tpp, coarse_pass=False)
    o.find_mu0()
    mu_fin = o.s.mu0
    mu_inf = post.mu0_infinite(n, T, tp, tpp, 0.0)
    assert abs(mu_inf - mu_fin) < 1e-10
    assert abs(post.density_infinite(mu_inf, T, tp, tpp, 0.0) - n) < 1e-13
    # a coarse lattice is measurably off, so the comparison above is not vacuous
    o16 = oracle_mod.Oracle(8, 32, n, T, 2.0, tp=tp, tpp=tpp, coarse_pass=False)
    o16.find_mu0()
    assert abs(o16.s.mu0 - mu_inf) > 1e-6


def test_fermi_surface_at_half_filling_is_the_square():
    fs = post.fermi_surface(0.0, 0.0, 0.0, 0.0, nk=101)
    kx0 = np.arange(len(fs)) * PI / 100
    assert len(fs) == 101
    assert np.allclose(fs[:, 0], 0.5 * (PI + kx0), atol=1e-13, rtol=0)
    assert np.allclose(fs[:, 1], 0.5 * (PI - kx0), atol=1e-13, rtol=0)


def test_fermi_surface_lies_on_the_level_set():
    tp, tpp, mu0 = -0.24, 0.15, -0.6
    fs = post.fermi_surface(mu0, tp, tpp, 0.0, nk=201)
    assert 0 < len(fs) <= 201
    e = post.dispersion(fs[:, 0], fs[:, 1], tp, tpp, 0.0)
    assert np.abs(e - mu0).max() < 1e-12
    kx0 = np.arange(len(fs)) * PI / 200
    assert np.all(fs[:, 0] >= kx0 - 1e-15) and np.all(fs[:, 0] <= PI + 1e-15)


def test_lattice_fermi_surface_is_the_minimum_of_abs_xi_on_each_diagonal():
    nbk = 9
    k = np.arange(nbk) * PI / (nbk - 1)
    ek = post.dispersion(k[None, :], k[:, None], -0.2, 0.05, 0.0).ravel()  # ek[x + y*nbk]
    mu0 = -0.3
    idx, amax = post.fs_lattice(ek, nbk, mu0)
    for l in range(nbk):
        xs = np.array([abs(ek[(l + m) + m * nbk] - mu0) for m in range(nbk - l)])
        assert idx[l] == int(np.argmin(xs))
        eta = 1e-5
        assert abs(amax[l] - eta / (PI * (eta * eta + xs.min() ** 2))) < 1e-9 * amax[l]


def _poles(k):
    """a two-pole self-energy per wedge point: Sigma(z) = sum_p a_p/(z - e_p) (rational, degree 1/2)"""
    a = np.array([0.7 + 0.01 * k, 0.4])
    e = np.array([-1.1 + 0.02 * k, 0.9 - 0.01 * k])
    return a, e


def _sigma(k, z):
    a, e = _poles(k)
    return np.sum(a[:, None] / (np.atleast_1d(z)[None, :] - e[:, None]), axis=0)


@pytest.mark.parametrize("use_pade", [True, False])
def test_mdc_recovers_a_rational_self_energy(use_pade):
    nbk, T, mu = 5, 0.1, 0.2
    k1 = np.arange(nbk) * PI / (nbk - 1)
    ek = post.dispersion(k1[:, None], k1[None, :], -0.1, 0.0, 0.0).ravel()  # ek[m + l*nbk]
    nidx = np.arange(8)
    z0 = 1j * (2 * nidx + 1) * PI * T
    if use_pade:
        def sig_of(j, k):
            return _sigma(k, z0[j])[0]
    else:
        # a quadratic in z is reproduced by the Lagrange polynomial through 4 points
        def sig_of(j, k):
            z = z0[j]
            return (0.3 + 0.01 * k) + (0.2 - 0.1j) * z + 0.05 * z * z
    w, etas = 0.15, [0.05, 0.01]
    selfw, akw = post.mdc(ek, nbk, mu, sig_of, nidx, T, w, etas, use_pade=use_pade)
    for l in range(nbk):
        for m in range(l + 1):
            k = post.widx(l, m)
            for e, eta in enumerate(etas):
                z = w + 1j * eta
                exact = _sigma(k, z)[0] if use_pade else (0.3 + 0.01 * k) + (0.2 - 0.1j) * z + 0.05 * z * z
                assert abs(selfw[k, e] - exact) < 1e-10 * max(1, abs(exact))
                ekmu = ek[m + l * nbk] - mu
                a = -(1 / (z - ekmu - exact)).imag / PI
                assert abs(akw[k, e] - a) < 1e-9 * max(1, abs(a))


def test_non_uniform_frequency_set():
    ind = post.non_uniform_freqs(64, 2)
    # N1 = 8 consecutive, then two groups of N2 = 4 with steps 2 and 4
    assert list(ind) == [0, 1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28]
    assert list(post.non_uniform_freqs(64, 0)) == list(range(32))


def test_edc_recovers_a_rational_self_energy():
    nbk, T, mu, nbw = 5, 0.1, 0.1, 64
    k1 = np.arange(nbk) * PI / (nbk - 1)
    ek = post.dispersion(k1[:, None], k1[None, :], -0.1, 0.0, 0.0).ravel()  # ek[ky + kx*nbk]
    z0 = 1j * (2 * np.arange(nbw // 2) + 1) * PI * T

    def sig_of(j, k):
        return _sigma(k, z0[j])[0]

    w = np.linspace(-2, 2, 9)
    etas = [0.05]
    kxs, kys = [1, 3, -1, 6], [0, 2, 2, 1]  # negative and > nbk-1 indices fold back
    out = post.edc(ek, nbk, mu, sig_of, T, nbw, kxs, kys, w, 0, 8, etas)
    Nk = 2 * (nbk - 1)
    for p, (kx, ky) in enumerate(zip(kxs, kys)):
        kx, ky = kx % Nk, ky % Nk
        kx, ky = (Nk - kx if kx > nbk - 1 else kx), (Nk - ky if ky > nbk - 1 else ky)
        x, y = max(kx, ky), min(kx, ky)
        k = post.widx(x, y)
        ekmu = ek[ky + kx * nbk] - mu
        for i, wi in enumerate(w):
            s = _sigma(k, wi + 0.05j)[0]
            a = -(1 / ((wi - ekmu) + 0.05j - s)).imag / PI
            assert abs(out[p, i, 0] - a) < 1e-9 * max(1, a)


def test_self_real_freq_recovers_rational_self_energies_in_both_forms():
    """Sigma = sum_p a_p/(z - e_p): Sinf = sum a_p, C2 = sum a_p e_p, C3 = sum a_p e_p^2 and
    R = sum a_p e_p^3 z/(z - e_p).  Four Pade points reproduce two poles (a [1/2] rational); in the
    continued-fraction form Sigma = Sinf/(z + b1 + S(z)) of three poles, S is a [1/2] rational too,
    so its four-point Pade fit is exact."""
    nbk, T, mu = 3, 0.1, 0.05
    nk8 = nbk * (nbk + 1) // 2
    ek = np.linspace(-1, 1, nbk * nbk)
    nidx = np.arange(4)
    z0 = 1j * (2 * nidx + 1) * PI * T
    w = np.array([-0.5, 0.0, 0.7])
    etas = [0.02, 0.3]
    for a, e, form in ((np.array([0.8, 0.3]), np.array([-0.35, 0.6]), 0),
                       (np.array([0.8, 0.3, 0.5]), np.array([-0.35, 0.6, 1.4]), 1)):
        C2 = np.full(nk8, np.sum(a * e))
        C3 = np.full(nk8, np.sum(a * e * e))

        def sig_of(j, k):
            return np.sum(a / (z0[j] - e))

        def R_of(j, k):
            return np.sum(a * e ** 3 * z0[j] / (z0[j] - e))

        sw, ak, sj, aj = post.self_real_freq(ek, nbk, mu, np.sum(a), C2, C3, R_of, T, (1, 2), nidx, w, etas)
        ekmu = ek[2 + 1 * nbk] - mu
        for i, wi in enumerate(w):
            for q, eta in enumerate(etas):
                z = wi + 1j * eta
                exact = np.sum(a / (z - e))
                got = (sw, sj)[form][i, q]
                assert abs(got - exact) < 1e-11 * max(1, abs(exact))
                A = -((1 / PI) / (z - ekmu - exact)).imag
                assert abs((ak, aj)[form][i, q] - A) < 1e-10 * max(1, abs(A))


def test_fs_inter_polynomial_is_exact_for_a_quadratic_self_energy():
    nbk, T, mu = 6, 0.05, -0.2
    nk8 = nbk * (nbk + 1) // 2
    k1 = np.arange(nbk) * PI / (nbk - 1)
    ek = post.dispersion(k1[:, None], k1[None, :], -0.15, 0.0, 0.0).ravel()
    z0 = 1j * (2 * np.arange(4) + 1) * PI * T
    coef = [(0.1 + 0.01 * k, -0.3j, 0.02) for k in range(nk8)]
    sig = np.array([[c[0] + c[1] * z + c[2] * z * z for c in coef] for z in z0])
    (ip, ap), _ = post.fs_inter(ek, nbk, mu, sig, T)
    z = 1e-6j
    for l in range(nbk):
        vals = []
        for m in range(nbk - l):
            x, y = l + m, m
            c = coef[post.widx(x, y)]
            s = c[0] + c[1] * z + c[2] * z * z
            vals.append(-(1 / (z - (ek[x + y * nbk] - mu) - s)).imag / PI)
        assert ip[l] == int(np.argmax(vals))
        assert abs(ap[l] - max(vals)) < 1e-9 * max(vals)
