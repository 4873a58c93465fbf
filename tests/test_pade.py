"""Host Pade continuation of the library (generique.cpp:1605-1663, chi.cpp:1055-1093) on the
CPU: the same algorithm as the oracle's restatement, exact interpolation, and recovery of
rational functions whose spectra are known in closed form."""
import numpy as np
import pytest

from conftest import relerr


@pytest.fixture(scope="module")
def capi(hip_lib):
    from tpsc_conductivity_amd import capi

    return capi


def test_coefficients_and_values_match_the_oracle(capi, oracle_mod):
    rng = np.random.default_rng(7)
    T = 0.1
    z = 1j * 2 * np.pi * T * np.arange(12)
    f = rng.normal(size=12) + 1j * rng.normal(size=12)
    c_lib = capi.pade_coef(z, f)
    x = np.linspace(0.0, 4.0, 9) + 0.05j
    v_lib = capi.pade_eval(z, c_lib, x)
    v_orc = oracle_mod.pade(f, z, x)
    assert relerr(v_lib, v_orc) < 1e-12


def test_approximant_interpolates_its_data(capi):
    rng = np.random.default_rng(11)
    z = 1j * (2 * np.arange(10) + 1) * np.pi * 0.2
    f = rng.normal(size=10) + 1j * rng.normal(size=10)
    c = capi.pade_coef(z, f)
    assert relerr(capi.pade_eval(z, c, z), f) < 1e-10


def test_rational_function_is_recovered_off_the_axis(capi):
    """f(z) = sum_k a_k/(z - p_k) with 3 poles: 2*3+1 Matsubara points fix it exactly."""
    a = np.array([0.7, 0.2 - 0.1j, 0.5 + 0.3j])
    p = np.array([1.3 - 0.4j, -0.8 - 0.2j, 2.5 - 1.0j])
    f = lambda z: np.sum(a[None, :] / (np.atleast_1d(z)[:, None] - p[None, :]), axis=1)
    z = 1j * 2 * np.pi * 0.1 * np.arange(1, 8)
    c = capi.pade_coef(z, f(z))
    x = np.linspace(-3, 3, 25) + 0.1j
    assert relerr(capi.pade_eval(z, c, x), f(x)) < 1e-8


def test_conductivity_of_a_known_bosonic_spectrum(capi):
    """Lambda(z) = sum_k A_k (1/(z - W_k) - 1/(z + W_k)): real and even on iq_n; on the real axis
    Im Lambda(w + i eta) is a pair of Lorentzians per mode, so sigma(w) is known in closed form."""
    A = np.array([0.3, 0.9])
    W = np.array([0.8, 2.1])
    T = 0.05
    lam_z = lambda z: np.sum(A[None, :] * (1 / (z[:, None] - W[None, :]) - 1 / (z[:, None] + W[None, :])), axis=1)
    q = 2 * np.pi * T * np.arange(12)
    lam = lam_z(1j * q).real
    w = np.array([0.0, 0.3, 0.8, 1.5, 2.1, 3.0])
    eta = 0.05
    w_eff = np.where(w == 0.0, 1e-6, w)
    exact = lam_z(w_eff + 1j * eta).imag / w_eff
    sig = capi.pade_conductivity(lam, T, 12, w, eta)
    assert relerr(sig, exact) < 1e-7


def test_bad_arguments_are_rejected(capi):
    with pytest.raises(capi.TPSCError):
        capi.pade_conductivity(np.ones(4), 0.1, 8, [1.0], 0.1)  # npade > len(lam)
    with pytest.raises(capi.TPSCError):
        capi.pade_conductivity(np.ones(4), -0.1, 4, [1.0], 0.1)  # T <= 0
