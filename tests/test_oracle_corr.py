"""Pins of the oracle's spin-correlation-length diagnostics (chi::sp_corr_length,
chi.cpp:451-640, and the hot-spot Fermi velocity, chi.cpp:419-447 / green0.cpp:626-675).

* vF: at half filling with t' = 0 the Fermi surface is the square kx + ky = pi and the hot spot
  (pi/2, pi/2) has |grad eps| = 2 sqrt 2; with t' != 0 the value equals |grad eps| at the root of
  eps(k) = mu0 on kx + ky = pi found directly, to the resolution of the 10001-point surface.
* the asymptotic length: for chi0 quadratic around (pi, pi) the discrete curvature is exact,
  xi_as = sqrt(a) sqrt(Usp/(2/chi0max - Usp)).
* the half-width length: for a Lorentzian chi_sp(q) = A/(1 + xi^2 |q - Q|^2) it returns xi up
  to the linear interpolation of the half-maximum point.
"""
import numpy as np
from scipy.optimize import brentq


def test_hot_spot_velocity(oracle_mod):
    o = oracle_mod.Oracle(16, 64, 1.0, 0.2, 2.0, coarse_pass=False)
    assert abs(oracle_mod.lib().orc_hot_spot_vF(o._o) - 2 * np.sqrt(2)) < 1e-7
    tp = -0.2
    o = oracle_mod.Oracle(16, 64, 0.9, 0.2, 2.0, tp=tp, coarse_pass=False)
    mu0 = o.s.mu0
    kx = brentq(lambda x: o.dispersion(x, np.pi - x) - mu0, 0.0, np.pi / 2)
    ky = np.pi - kx
    gx, gy = o.velocity_x(kx, ky), o.velocity_x(ky, kx)
    assert abs(oracle_mod.lib().orc_hot_spot_vF(o._o) - np.hypot(gx, gy)) < 5e-4 * np.hypot(gx, gy)


def wedge(f, nbq):
    d = np.pi / (nbq - 1)
    return np.array([f(l * d, m * d) for l in range(nbq) for m in range(l + 1)])


def test_asymptotic_length_of_a_quadratic_peak(oracle_mod):
    nbq, cmax, a, Usp = 65, 0.8, 2.5, 2.0
    c0 = wedge(lambda x, y: cmax * (1 - a * ((x - np.pi) ** 2 + (y - np.pi) ** 2)), nbq)
    out = oracle_mod.sp_corr_from(c0, nbq, nbq - 1, nbq - 1, cmax, Usp)
    want = np.sqrt(a) * np.sqrt(Usp / (2 / cmax - Usp))
    assert abs(out[2] - want) < 1e-12 * want and abs(out[3] - want) < 1e-12 * want


def test_half_width_length_of_a_lorentzian_peak(oracle_mod):
    nbq, A, xi, Usp = 129, 20.0, 4.0, 2.0
    csp = lambda x, y: A / (1 + xi ** 2 * ((x - np.pi) ** 2 + (y - np.pi) ** 2))
    c0 = wedge(lambda x, y: csp(x, y) / (1 + Usp * csp(x, y) / 2), nbq)
    cmax = c0[-1]
    out = oracle_mod.sp_corr_from(c0, nbq, nbq - 1, nbq - 1, cmax, Usp)
    assert abs(out[0] - xi) < 1e-2 * xi and out[1] == out[0]
