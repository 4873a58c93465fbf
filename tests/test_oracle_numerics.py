"""CPU tests pinning the oracle's numerical kernels to what the mathematics fixes.

The reference delegates these to FFTW3 (REDFT00/RODFT00/dft, src/CMakeLists.txt:11) and
LAPACK dgbsv_ (generique.cpp:303); neither ships golden vectors, so they are pinned here by
their definitions (scipy/numpy), uniqueness of the clamped spline, exact integrals of
piecewise cubics and known roots.
"""
import numpy as np
import pytest
from scipy.fft import dct, dst
from scipy.interpolate import CubicSpline

from conftest import relerr


@pytest.mark.parametrize("n0,n1", [(3, 3), (5, 9), (17, 17), (31, 33), (65, 65)])
def test_redft00_2d_matches_definition(oracle_mod, n0, n1):
    rng = np.random.default_rng(n0 * 100 + n1)
    a = rng.standard_normal((n0, n1))
    ref = dct(dct(a, type=1, axis=1), type=1, axis=0)
    assert relerr(oracle_mod.r2r_2d(a, 0, 0), ref) < 1e-14


@pytest.mark.parametrize("n0,n1", [(1, 3), (3, 5), (15, 17), (63, 65)])
def test_rodft00_x_redft00_matches_definition(oracle_mod, n0, n1):
    rng = np.random.default_rng(n0 + 7 * n1)
    a = rng.standard_normal((n0, n1))
    ref = dct(dst(a, type=1, axis=0), type=1, axis=1)
    assert relerr(oracle_mod.r2r_2d(a, 1, 0), ref) < 1e-14


def test_redft00_brute_force_small(oracle_mod):
    # FFTW REDFT00 written out: Y_k = X0 + (-1)^k X_{n-1} + 2 sum_j X_j cos(pi j k/(n-1))
    x = np.array([0.3, -1.2, 2.5, 0.7, -0.1])
    n = len(x)
    y = [x[0] + (-1) ** k * x[-1] + 2 * sum(x[j] * np.cos(np.pi * j * k / (n - 1)) for j in range(1, n - 1))
         for k in range(n)]
    assert relerr(oracle_mod.redft00(x), y) < 1e-15
    for m in (3, 9, 33, 1025):
        v = np.random.default_rng(m).standard_normal(m)
        assert relerr(oracle_mod.redft00(v), dct(v, type=1)) < 1e-13


@pytest.mark.parametrize("n", [8, 64, 512, 2048, 6, 12])
def test_dft_both_signs(oracle_mod, n):
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert relerr(oracle_mod.dft(z, -1), np.fft.fft(z)) < 1e-14
    assert relerr(oracle_mod.dft(z, +1), np.fft.ifft(z) * n) < 1e-14


@pytest.mark.parametrize("N0,T", [(5, 1.0), (41, 1.0), (257, 0.2), (2049, 0.05)])
def test_spline_is_the_clamped_cubic_spline(oracle_mod, N0, T):
    x = np.arange(N0) / ((N0 - 1) * T)
    y = np.sin(3 * x) + 0.1 * x * x * np.cos(x)
    d0, dN = 1.3, -0.7
    c = oracle_mod.spline_coeffs_rel(x, y, d0, dN)
    ref = CubicSpline(x, y, bc_type=((1, d0), (1, dN))).c.T
    for col in range(4):
        assert relerr(c[:, col], ref[:, col]) < 5e-13


def test_spline_nonuniform_grid(oracle_mod):
    x = np.cumsum(np.r_[0.0, np.random.default_rng(3).uniform(0.05, 0.2, 30)])
    y = np.exp(-x) * np.cos(4 * x)
    c = oracle_mod.spline_coeffs_rel(x, y, -1.0, 0.2)
    ref = CubicSpline(x, y, bc_type=((1, -1.0), (1, 0.2))).c.T
    assert relerr(c, ref) < 1e-13


def test_spline_reproduces_cubics_exactly(oracle_mod):
    # a cubic with its own end slopes is its own clamped spline (uniqueness)
    x = np.linspace(0, 3, 25)
    p = lambda t: 0.5 * t ** 3 - 2 * t ** 2 + t - 4
    dp = lambda t: 1.5 * t ** 2 - 4 * t + 1
    c = oracle_mod.spline_coeffs_rel(x, p(x), dp(x[0]), dp(x[-1]))
    assert np.allclose(c[:, 0], 0.5, rtol=0, atol=1e-11)


def test_spline_fourier_formula_equals_exact_integral(oracle_mod):
    """chi0.cpp:610-626 / green.cpp:1955-2006: int_0^beta e^{i w_n tau} S(tau) dtau of a
    clamped spline equals the 1/z..1/z^4 moments plus the DFT of 6a_j (fermionic case)."""
    T, Nw = 0.5, 64
    beta = 1 / T
    tau = np.arange(Nw + 1) / (Nw * T)
    y = np.cos(2 * tau) * np.exp(-0.3 * tau) + 0.2 * tau
    dyf = lambda t: -2 * np.sin(2 * t) * np.exp(-0.3 * t) - 0.3 * np.cos(2 * t) * np.exp(-0.3 * t) + 0.2
    d0, dN = dyf(0.0), dyf(beta)
    c = oracle_mod.spline_coeffs_rel(tau, y, d0, dN)
    h = tau[1] - tau[0]
    X = np.exp(1j * np.pi * np.arange(Nw) / Nw) * 6 * c[:, 0]
    F = oracle_mod.dft(X, +1)
    S0, Sb = y[0], y[-1]
    d2S0 = 2 * c[0, 1]
    d2Sb = 6 * c[-1, 0] * h + 2 * c[-1, 1]
    for n in range(0, 6):
        wn = (2 * n + 1) * np.pi * T
        z = 1j * wn
        R = (1 - np.exp(1j * np.pi * (2 * n + 1) / Nw)) * F[n]
        formula = -(S0 + Sb) / z + (d0 + dN) / z ** 2 - (d2S0 + d2Sb) / z ** 3 + R / z ** 4
        # exact integral of the piecewise cubic, interval by interval (Gauss-Legendre, exact)
        xg, wg = np.polynomial.legendre.leggauss(8)
        tot = 0j
        for j in range(Nw):
            t = tau[j] + (xg + 1) * h / 2
            s = c[j, 0] * (t - tau[j]) ** 3 + c[j, 1] * (t - tau[j]) ** 2 + c[j, 2] * (t - tau[j]) + c[j, 3]
            tot += np.sum(wg * s * np.exp(1j * wn * t)) * h / 2
        assert abs(formula - tot) < 1e-12 * max(1.0, abs(tot))


def test_find_zero_bracketed_and_limits(oracle_mod):
    f = lambda x: x ** 3 - 2 * x - 5  # root 2.0945514815423265
    ok, r, ne = oracle_mod.find_zero(f, [1.0, 0.0, 4.0])
    assert ok and abs(r - 2.0945514815423265) < 1e-14 and ne <= 4 + 2 * 20
    ok, r, _ = oracle_mod.find_zero(f, [1.0, 0.0, 4.0], lims=[-20, 20])
    assert ok and abs(r - 2.0945514815423265) < 1e-13
    # no sign change in the interval: secant search leaves it (generique.cpp:640-723)
    g = lambda x: (x - 5.5) * (1.0 + 0.1 * x * x)
    ok, r, _ = oracle_mod.find_zero(g, [0.0, 1.0, 3.0])
    assert ok and abs(r - 5.5) < 1e-12
    # with limits the escaped bracket is rejected (generique.cpp:1040-1044)
    ok, _, _ = oracle_mod.find_zero(g, [0.0, 1.0, 3.0], lims=[-5, 5])
    assert not ok
    # a flat function defeats the secant search and is reported as failure
    ok, _, _ = oracle_mod.find_zero(lambda x: np.tanh(x - 7.0), [0.0, 1.0, 3.0])
    assert not ok


def test_roots_2pol(oracle_mod):
    for coef in ([1, -3 + 0.5j, 2 - 1j], [1, 0.2, -4.3], [2, 1j, 1]):
        r = oracle_mod.roots_2pol(coef)
        ref = np.roots(coef)
        for z in r:
            assert np.min(np.abs(ref - z)) < 1e-14 * max(1.0, abs(z))
        for z in ref:
            assert np.min(np.abs(r - z)) < 1e-14 * max(1.0, abs(z))
        # polished roots satisfy the polynomial to rounding
        assert np.max(np.abs(np.polyval(coef, r))) < 1e-14


def test_pade_interpolates_and_recovers_rational(oracle_mod):
    # Thiele continued fraction through N points reproduces them; a [1/2] rational
    # function is recovered exactly from 4 points.
    x = 1j * np.pi * 0.1 * (2 * np.arange(6) + 1)
    f = lambda z: (z + 0.3) / ((z - 0.5 + 0.1j) * (z + 1.2))
    vals = f(x)
    out = oracle_mod.pade(vals, x, x)
    assert relerr(out, vals) < 1e-12
    zs = np.array([0.3 + 0.05j, -0.7 + 0.2j, 2.0 + 0.01j])
    assert relerr(oracle_mod.pade(vals, x, zs), f(zs)) < 1e-9


@pytest.mark.parametrize("N,Nw,T", [(16, 256, 0.1), (32, 1024, 0.05)])
def test_self_energy_remainder_is_the_exact_transform_of_its_inputs(N, Nw, T, oracle_mod):
    """The raw remainder R(k, n) (green.cpp:1925-2018) equals its transform of Sigma(k, tau) and the
    end slopes recomputed in x87 extended precision (tests/remainder_ld.py) to ~1e-12: the spline
    solve and the phased DFT add no error of their own beyond FP64 rounding."""
    import remainder_ld as rl

    o = oracle_mod.Oracle(N, Nw, 0.875, T, 4.0, tp=-0.175, coarse_pass=False)
    assert o.solve()
    nk8 = o.nk8
    St = o.get("Sigma_tau").reshape(Nw + 1, nk8)
    dS = o.get("dtau_Sigma").reshape(2, nk8)
    R = o.get("Self").reshape(-1, nk8)[: Nw // 2]
    ks = np.arange(nk8) if nk8 < 60 else np.linspace(0, nk8 - 1, 8).astype(int)
    Rl = rl.remainder(St[:, ks], dS[0, ks], dS[1, ks], T).astype(np.complex128)
    e = np.max(np.abs(R[:, ks] - Rl), axis=0) / np.max(np.abs(Rl), axis=0)
    assert e.max() < 1e-11
