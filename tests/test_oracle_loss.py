"""The oracle's loss terms (orc_loss), the reference side of the gradient tests, against the
quantities they are built from: the RE term is orc_sigma_loss bit for bit; the IM term is the
moment form of Sigma(k, i pi T) (pinned against a brute-force Matsubara sum in
test_oracle_physics.py) summed in numpy; the chi_sp term sits at q = (pi, pi), where chi0(q, 0)
peaks at half filling with t' = 0; the Lambda_jj term is the sum over n <= 8 of the arrays of a
plain solve; and the terms add up."""
import numpy as np
import pytest

RE, IM, CHISP, LAM = 1, 2, 4, 8
PI = np.pi


def make(oracle_mod, n=0.9, tp=-0.15, U=3.0):
    return oracle_mod.Oracle(16, 64, n, 0.2, U, tp=tp, coarse_pass=False)


@pytest.fixture(scope="module")
def w(oracle_mod):
    o = make(oracle_mod)
    return np.linspace(0.1, 1.0, o.nk8)


def test_re_term_is_the_sigma_loss(oracle_mod, w):
    assert make(oracle_mod).loss(w, RE) == make(oracle_mod).sigma_loss(w)


def test_im_term_from_the_moment_form(oracle_mod, w):
    o = make(oracle_mod)
    L = o.loss(w, IM)
    s = o.scalars()
    nk8 = o.nk8
    R0 = o.get("Self").reshape(-1, nk8)[0]
    z = 1j * PI * 0.2
    S = s["Sigma_inf"] / z + o.get("Sigma_inf2") / z ** 2 + o.get("Sigma_inf3") / z ** 3 + R0 / z ** 4
    assert abs(L - np.sum(w * S.imag)) <= 1e-13 * abs(L)
    assert L < 0  # Im Sigma(k, i w_0) < 0 for w_0 > 0


def test_chisp_term_at_the_antiferromagnetic_wave_vector(oracle_mod):
    o = make(oracle_mod, n=1.0, tp=0.0, U=2.0)
    L = o.loss(None, CHISP)
    s = o.scalars()
    c = o.get("chiqw")[o.nk8 - 1]
    assert c == s["chi0max"]  # perfect nesting: chi0(q, 0) is largest at q = (pi, pi)
    assert abs(L - c / (1 - s["Usp"] * c / 2)) <= 1e-15 * abs(L)
    assert L > c > 0  # enhanced by the spin vertex


def test_lambda_term_and_the_sum_of_terms(oracle_mod, w):
    o = make(oracle_mod)
    assert o.solve()
    lam = o.get("chijj1") + o.get("chijj2")
    a = make(oracle_mod).loss(w, LAM)
    assert a == np.sum(lam[:9]) or abs(a - np.sum(lam[:9])) <= 1e-15 * abs(a)
    parts = [make(oracle_mod).loss(w, b) for b in (RE, IM, CHISP, LAM)]
    tot = make(oracle_mod).loss(w, RE | IM | CHISP | LAM)
    assert abs(tot - sum(parts)) <= 1e-13 * sum(abs(p) for p in parts)
