"""Pins of the oracle's Aslamazov-Larkin restatement (cond_opt.cpp:4292-6467) by what the
mathematics fixes, independently of the reference's code path.

* hG(q,tau) = DST x DCT [h(r,tau) G0(r,-tau)] is the vertex-weighted particle-hole bubble
  (1/N) sum_k v_k G0(k,tau) G0(k+q,-tau) on the N x N lattice, so
    hg2(q)     = (2/N) sum_k v_k (f_k - f_k+q) (e_k+q - e_k)           (exact: slopes at 0, beta)
    hG(q,iq_p) = -(2/N) sum_k v_k (f_k - f_k+q) D/(D^2 + q_p^2),  D = e_k+q - e_k
  (generalized Lindhard sum); the oracle's spline transform must converge to it as O(1/Nw^4).
* The imaginary-time two-pole tail equals the direct Matsubara sum of its frequency form;
  its tau = beta value and its end slopes are the limits / derivatives of the general form.
* The sparse boson set follows the reference's rule; the angle bins add up to the total.
* The k-space, fermionic-frequency stage (q -> r, products with G0(r,-tau), r -> k, splines,
  shifted transforms, the G G sum) converges to the direct Matsubara-convolution form of the
  same expression (al_brute) as Nw grows.
"""
import numpy as np
import pytest
from scipy.signal import fftconvolve


def lattice(o, N):
    ks = 2 * np.pi * np.arange(N) / N
    e = np.array([[o.dispersion(a, b) for b in ks] for a in ks])
    v = np.array([[o.velocity_x(a, b) for b in ks] for a in ks])
    return e, v


def lindhard(o, N, T, mu0, p_list):
    """hg2 and hG(q,iq_p) on the Lambda grid q = (l+1, m) by direct lattice sums."""
    e, v = lattice(o, N)
    f = 1.0 / (np.exp((e - mu0) / T) + 1.0)
    nk = N // 2 + 1
    hg2 = np.zeros(nk * (nk - 2))
    hw = np.zeros((len(p_list), nk * (nk - 2)))
    for l in range(nk - 2):
        for m in range(nk):
            q = m + l * nk
            eq = np.roll(np.roll(e, -(l + 1), 0), -m, 1)
            fq = np.roll(np.roll(f, -(l + 1), 0), -m, 1)
            D = eq - e
            hg2[q] = 2.0 / N ** 2 * np.sum(v * (f - fq) * D)
            deg = np.abs(D) < 1e-12
            for i, p in enumerate(p_list):
                w = 2 * np.pi * p * T
                if p == 0:
                    r = np.where(deg, f * (1 - f) / T, (f - fq) / np.where(deg, 1.0, D))
                else:
                    r = (f - fq) * D / (D * D + w * w)
                hw[i, q] = -2.0 / N ** 2 * np.sum(v * r)
    return hg2, hw


def al_brute(o, n_idx, P):
    """Lambda_jj3 at the n_idx-th sparse frequency by direct Matsubara convolutions:
    HG(k,iw) = -(T/N^2) sum_{q,p} HG(q,iq_p) / (i(w + q_p) - xi_k+q),
    HG2(k,iw) = -(T/N^2) sum_{q,p} HG(q,iq_p) / (i(w - q_p) - xi_k+q),  |p| <= P,
    i.e. the k-space, fermionic-frequency form of [HG(r,tau) G0(r,-tau)] and
    [HG(r,beta-tau) G0(r,-tau)] (DST x DCT = -i x lattice Fourier sum for odd-in-x functions),
    with HG(q,iq_p) from the oracle's chi0(q,iq_p), hG(q,iq_p), hg2 and tails on the full
    lattice (odd in qx).  No tau grid, no splines, no two-pole tail model."""
    s = o.s; N = s.p.nk0; nbw = s.nbw; H = nbw // 2; T = s.p.T; mu0 = s.mu0; mu = s.mu
    nk = N // 2 + 1; nk8 = nk * (nk + 1) // 2; nq = nk * (nk - 2); Nsites = N * N
    Usp, Uch, U = s.Usp, s.Uch, s.p.U
    chiqw = o.get("chiqw"); ci = o.get("chi0_inf"); ci2 = o.get("chi0_inf2")
    hw = o.get("al_hw").reshape(H, nq); hg2 = o.get("al_hg2")
    Self = o.get("Self"); S2 = o.get("Sigma_inf2"); S3 = o.get("Sigma_inf3"); Sinf = s.Sigma_inf
    nn = int(o.al_freqs()[n_idx]); nc = nn // 2; qn = 2 * np.pi * nn * T
    def widx(l, m): return m + l * (l + 1) // 2
    def fold(x, y):
        if x > N // 2: x = N - x
        if y > N // 2: y = N - y
        if y > x: x, y = y, x
        return widx(x, y)
    ks = 2 * np.pi * np.arange(N) / N
    eps = np.array([[o.dispersion(a, b) for b in ks] for a in ks])
    p = np.arange(-P, P + 1)
    ap = np.abs(p); w = 2 * np.pi * ap * T
    # A(q_full, p)
    A = np.zeros((N, N, 2 * P + 1))
    for qx in range(N):
        if qx == 0 or qx == N // 2: continue
        sg = 1.0 if qx < N // 2 else -1.0
        lx = qx if qx < N // 2 else N - qx
        for qy in range(N):
            ly = qy if qy <= N // 2 else N - qy
            qro = ly + (lx - 1) * nk
            r = fold(lx, ly)
            def chi0(pp):
                a = np.abs(pp); ww = 2 * np.pi * a * T
                out = np.empty(len(a))
                m = a <= H
                out[m] = chiqw[r + a[m] * nk8]
                ww2 = ww[~m]
                out[~m] = ci[r] / ww2**2 + ci2[r] / ww2**4
                return out
            def hG(pp):
                a = np.abs(pp); ww = 2 * np.pi * a * T
                out = np.empty(len(a))
                z = a == 0
                out[z] = hw[0, qro]
                nz = ~z
                c4 = np.where(a[nz] < H, hw[np.minimum(a[nz], H - 1), qro], hw[H - 1, qro])
                out[nz] = -hg2[qro] / ww[nz]**2 + c4 / ww[nz]**4
                return out
            pe = np.arange(-P - 2, P + 1)
            with np.errstate(divide='ignore', invalid='ignore'):
                D = (hG(pe) - hG(pe + nn)) / (chi0(pe) - chi0(pe + nn))
            Dp = D[2:]
            if nn % 2 == 0:
                i = -nn // 2 + P
                Dp[i] = (4 * D[i + 1] - D[i]) / 3.0   # D[i+2-1], D[i+2-2]
            c1 = chi0(p)
            G = 3 * Usp * c1 / (1 - Usp * c1 / 2) + Uch * c1 / (1 + Uch * c1 / 2)
            A[qx, qy] = sg * G * Dp
    A = A.reshape(N * N, 2 * P + 1)
    jj = np.arange(nbw); f = np.where(jj < H - nc, jj, jj - nbw)
    smin, smax = f.min(), f.max() + nn
    umin, umax = smin - P, smax + P
    u = np.arange(umin, umax + 1)
    tot = 0.0
    for l in range(nk - 2):
        for m in range(nk):
            kx, ky = l + 1, m
            xi = np.array([eps[(kx + qx) % N, (ky + qy) % N] - mu0 for qx in range(N) for qy in range(N)])
            b = 1.0 / (1j * (2 * u[None, :] + 1) * np.pi * T - xi[:, None])
            c1 = fftconvolve(b, A[:, ::-1], axes=1).sum(0)
            c2 = fftconvolve(b, A, axes=1).sum(0)
            HGk = -T / Nsites * c1[(f + nn) + P - umin]
            HG2k = -T / Nsites * c2[f + P - umin]
            ki = fold(kx, ky); e = eps[kx, ky]
            def sigma(ff):
                kn = (2 * ff + 1) * np.pi * T
                g = np.where(ff >= 0, ff, -ff - 1)
                R = np.where(g < H, Self[ki + np.minimum(g, H - 1) * nk8], 0)
                R = np.where(ff >= 0, R, np.conj(R))
                return (-S2[ki] / kn**2 + R.real / kn**4) + 1j * (-Sinf / kn + S3[ki] / kn**3 + R.imag / kn**4)
            k1 = (2 * f + 1) * np.pi * T; k2 = (2 * (f + nn) + 1) * np.pi * T
            G1 = 1 / (1j * k1 - e + mu - sigma(f)); G2 = 1 / (1j * k2 - e + mu - sigma(f + nn))
            wt = 2.0 if (m == 0 or m == nk - 1) else 4.0
            v = o.velocity_x(kx * np.pi / (nk - 1), ky * np.pi / (nk - 1))
            tot += np.real(-1j * U * T / Nsites * wt * v * np.sum(G1 * G2 * (HGk - HG2k)) / (2 * qn))
    return tot


def solved(oracle_mod, N, Nw, mmin=2, T=0.1):
    o = oracle_mod.Oracle(N, Nw, 0.9, T, 2.0, tp=-0.2, tpp=0.05, coarse_pass=False)
    assert o.solve()
    assert o.calc_chijj_al(mmin, 1) == 0
    return o


def test_al_hg2_exact(oracle_mod):
    o = solved(oracle_mod, 8, 64)
    hg2, _ = lindhard(o, 8, o.s.p.T, o.s.mu0, [])
    got = o.get("al_hg2")
    assert np.max(np.abs(got - hg2)) <= 1e-12 * np.max(np.abs(hg2))


def test_al_hG_converges_to_lindhard(oracle_mod):
    errs = []
    for Nw in (64, 128, 256):
        o = solved(oracle_mod, 8, Nw)
        T = o.s.p.T
        p_list = [0, 1, 2, 5]
        hg2, ref = lindhard(o, 8, T, o.s.mu0, p_list)
        hw = o.get("al_hw").reshape(Nw // 2, o.nq)
        got = [hw[0]] + [-o.get("al_hg2") / (2 * np.pi * p * T) ** 2 + hw[p] / (2 * np.pi * p * T) ** 4
                         for p in p_list[1:]]
        errs.append(max(np.max(np.abs(got[i] - ref[i])) / np.max(np.abs(ref[i])) for i in range(len(p_list))))
    # O(h^4): each doubling of Nw divides the error by ~16
    assert errs[-1] < 1e-6
    assert errs[1] < errs[0] / 10 and errs[2] < errs[1] / 10


def test_al_tail_matsubara_sum(oracle_mod):
    Usp, Uch, ci, T = 2.3, 4.1, 0.7, 0.1
    P = 400000
    w = 2 * np.pi * np.arange(-P, P + 1) * T
    Hinf = 3 * Usp / (-w * w + Usp * ci / 2) + Uch / (-w * w - Uch * ci / 2)
    for tau in (0.3, 2.5, 7.0, 9.9):
        direct = T * np.sum(Hinf * np.exp(-1j * w * tau))
        assert abs(oracle_mod.al_tail(Usp, Uch, ci, T, tau, 0) - direct) < 1e-8
    beta = 1 / T
    vb = oracle_mod.al_tail(Usp, Uch, ci, T, beta, 1)
    assert abs(vb - oracle_mod.al_tail(Usp, Uch, ci, T, beta * (1 - 1e-13), 0)) < 1e-10
    h = 1e-5
    for kind, t0, sgn in ((2, 0.0, 1.0), (3, beta, -1.0)):
        v0 = oracle_mod.al_tail(Usp, Uch, ci, T, t0, 1 if kind == 3 else 0)
        v1 = oracle_mod.al_tail(Usp, Uch, ci, T, t0 + sgn * h, 0)
        v2 = oracle_mod.al_tail(Usp, Uch, ci, T, t0 + sgn * 2 * h, 0)
        fd = sgn * (-3 * v0 + 4 * v1 - v2) / (2 * h)
        assert abs(oracle_mod.al_tail(Usp, Uch, ci, T, t0, kind) - fd) < 1e-7 * max(1.0, abs(fd))


def test_al_frequency_set_rule(oracle_mod):
    # Nw = 128, m = 2: n = 0..16, then 8 steps of 2, then 8 steps of 4 up to N0 - 1 = 64
    f = oracle_mod.al_freqs(128, 2)
    assert list(f) == list(range(17)) + list(range(18, 33, 2)) + list(range(36, 65, 4))
    # the reference driver's grid (Nw = 4096, m = 3): 641 frequencies ending at Nw/2
    f = oracle_mod.al_freqs(4096, 3)
    assert len(f) == 641 and f[-1] == 2048 and np.all(np.diff(f) > 0)
    assert list(oracle_mod.al_freqs(64, 0)) == list(range(33))


def test_al_angle_bins_and_total(oracle_mod):
    o = oracle_mod.Oracle(16, 64, 0.9, 0.2, 2.0, tp=-0.2, coarse_pass=False)
    assert o.solve()
    assert o.calc_chijj_al(2, 4) == 0
    c3, k3 = o.get("chijj3"), o.get("chijj3_k").reshape(-1, 4)
    assert c3[0] == 0.0
    assert np.allclose(k3.sum(axis=1), c3, rtol=1e-12, atol=1e-15)
    f = o.al_freqs()
    tot = o.get("chijj1")[f] + o.get("chijj2")[f] + c3
    tot[0] = o.get("chijj1")[0] + o.get("chijj2")[0]
    assert np.allclose(o.get("chijj_all"), tot, rtol=1e-14)


def test_al_lambda3_converges_to_matsubara_convolution(oracle_mod):
    errs = {}
    for Nw in (64, 128):
        o = oracle_mod.Oracle(8, Nw, 0.9, 0.2, 2.0, tp=-0.2, tpp=0.05, coarse_pass=False)
        assert o.solve()
        assert o.calc_chijj_al(1, 1) == 0
        c3, fr = o.get("chijj3"), o.al_freqs()
        for n_idx in (1, 2, len(fr) - 1):  # odd n, even n (Richardson point), n = Nw/2
            b = al_brute(o, n_idx, 4096)
            errs[(Nw, n_idx if n_idx < 3 else -1)] = abs(c3[n_idx] - b) / abs(c3[n_idx])
    print("AL vs Matsubara convolution:", {k: f"{v:.1e}" for k, v in errs.items()})
    assert errs[(128, 1)] < 1e-6 and errs[(128, 2)] < 1e-6 and errs[(128, -1)] < 1e-3
    for n in (1, 2, -1):
        assert errs[(128, n)] < errs[(64, n)] / 3
