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
"""
import numpy as np
import pytest


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
