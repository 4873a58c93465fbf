"""The 1/z^4 remainder R(k, n) of Sigma(k, i w_n) recomputed in x87 extended precision (numpy
longdouble, 64-bit mantissa) from Sigma(k, tau_j), j = 0..L, and its end slopes: the clamped cubic
spline of the samples (moments M by the tridiagonal system of a uniform grid), then
R(k, n) = (1 - e^{i pi (2n+1)/L}) sum_j e^{i pi j/L} (M_{j+1} - M_j)/h e^{2 pi i j n/L}, n < L/2
(green.cpp:1925-2018; h = 1/(L T)).  Used by the tests of the remainder's precision."""
import numpy as np

LD = np.longdouble
CLD = np.clongdouble


def spline_moments(y, d0, dN, h):
    """Clamped-spline second derivatives of the columns of y ((L+1) x K) on a grid of step h with
    end slopes d0, dN (K each): [2 1; 1 4 1; ...; 1 2] M = r, by the Thomas algorithm."""
    L = y.shape[0] - 1
    r = np.empty_like(y)
    r[0] = 6 / h * ((y[1] - y[0]) / h - d0)
    r[1:L] = 6 / (h * h) * (y[2:] - 2 * y[1:L] + y[:L - 1])
    r[L] = 6 / h * (dN - (y[L] - y[L - 1]) / h)
    diag = np.full(L + 1, LD(4))
    diag[0] = diag[L] = LD(2)
    c = np.empty(L + 1, dtype=LD)
    d = np.empty_like(y)
    c[0] = LD(1) / diag[0]
    d[0] = r[0] / diag[0]
    for j in range(1, L + 1):
        den = diag[j] - c[j - 1]
        c[j] = LD(1) / den
        d[j] = (r[j] - d[j - 1]) / den
    M = np.empty_like(y)
    M[L] = d[L]
    for j in range(L - 1, -1, -1):
        M[j] = d[j] - c[j] * M[j + 1]
    return M


def remainder(y, d0, dN, T):
    y = np.asarray(y, dtype=LD)
    L = y.shape[0] - 1
    h = LD(1) / (LD(L) * LD(T))
    M = spline_moments(y, np.asarray(d0, dtype=LD), np.asarray(dN, dtype=LD), h)
    a6 = (M[1:] - M[:-1]) / h
    pi = LD("3.14159265358979323846264338327950288")
    j = np.arange(L)
    n = np.arange(L // 2)
    ph = np.exp(CLD(1j) * (pi * j.astype(LD) / LD(L)))
    ang = (2 * pi / LD(L)) * ((np.outer(n, j) % L).astype(LD))
    E = np.exp(CLD(1j) * ang)
    F = E @ (ph[:, None] * a6.astype(CLD))
    tw = 1 - np.exp(CLD(1j) * (pi * (2 * n + 1).astype(LD) / LD(L)))
    return tw[:, None] * F  # (L/2) x K
