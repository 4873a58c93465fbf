"""The split of the sum-rule passes (k_sum_rule3): rows whose elements all satisfy
|U chi/2| <= 1/16 enter as polynomials in y = -sU/2 over suffix moments M_m = sum w chi^m,
m = 2..21, the other rows directly.  Restated here in numpy on synthetic positive chi rows
that decay like 1/n^2 (as chi0(q, iqn) does), for both channels, U = 0 and an iterate large
enough that most rows are read directly; the three sums (trace, d/dU, d2/dU2 of
chi/(1 + sU chi/2) - chi, chi.cpp:738-830) must equal the direct sums to rounding."""
import numpy as np
import pytest

K, RHO = 18, 1.0 / 16.0


def _direct(c, w, sU, sgn):
    r = 1.0 / (1.0 + sU * c / 2.0)
    return np.array([np.sum(w * (c * r - c)), np.sum(-w * sgn * 0.5 * c * c * r * r),
                     np.sum(w * 0.5 * c ** 3 * r ** 3)])


def _split(c, w, sU, sgn):
    nrows = c.shape[0]
    R = np.array([[np.sum(w[j] * c[j] ** m) for m in range(2, K + 4)] for j in range(nrows)])
    SM = np.cumsum(R[::-1], axis=0)[::-1]
    smax = np.maximum.accumulate(np.abs(c).max(axis=1)[::-1])[::-1]
    Ux = abs(sU)
    thr = RHO / (0.5 * Ux) if Ux else np.inf
    n0 = int(np.sum(smax > thr))
    acc = _direct(c[:n0], w[:n0], sU, sgn)
    if n0 < nrows:
        M, y = SM[n0], -0.5 * sU
        f0 = f1 = f2 = 0.0
        for k in range(K, -1, -1):
            if k >= 1:
                f0 = f0 * y + M[k - 1]
            f1 = f1 * y + (k + 1) * M[k]
            f2 = f2 * y + 0.5 * (k + 1) * (k + 2) * M[k + 1]
        acc = acc + np.array([y * f0, -sgn * 0.5 * f1, 0.5 * f2])
    return acc, n0


@pytest.mark.parametrize("sgn,Ux", [(-1, 2.5), (-1, 0.3), (1, 7.0), (1, 0.0), (1, 40.0)])
def test_series_split_equals_direct_sums(sgn, Ux):
    rng = np.random.default_rng(7)
    nrows, nq = 400, 300
    n = np.arange(nrows)[:, None]
    c = rng.uniform(0.2, 0.6, size=(1, nq)) * 4.0 / (4.0 + (0.3 * n) ** 2)
    w = rng.uniform(0.5, 2.0, size=(1, nq)) * np.where(n == 0, 0.5, 1.0) * np.ones((nrows, 1))
    got, n0 = _split(c, w, sgn * Ux, sgn)
    ref = _direct(c, w, sgn * Ux, sgn)
    assert 0 <= n0 < nrows
    np.testing.assert_allclose(got, ref, rtol=2e-15, atol=0)


def test_series_truncation_bound():
    # the first omitted term of the c^3 r^3 series, relative to M_3: C(K+3, 2) rho^(K+1)
    assert (K + 2) * (K + 3) / 2 * RHO ** (K + 1) < 1e-20
