"""CPU tests pinning the oracle's TPSC chain.

The reference ships no tests or golden vectors; its runtime self-checks are the
known-answer checks (chi0.cpp:744-758, chi.cpp:177-179, green.cpp:2034-2052,3086-3089,
cond_opt.cpp:4282-4286).  They are turned into assertions here, next to closed forms and
independent restatements (brute-force Lindhard, brute-force Sigma, direct Matsubara sums
for the bubble, the density and kx0) chosen so that a dropped term, a wrong sign or weight
fails one of them.
"""
import numpy as np
import pytest

from conftest import relerr

PI = np.pi


def make(oracle_mod, N=16, Nw=256, n=0.875, T=0.1, U=4.0, tp=-0.175, tpp=0.0, t3=0.0, coarse=False):
    return oracle_mod.Oracle(N, Nw, n, T, U, tp=tp, tpp=tpp, t3=t3, coarse_pass=coarse)


@pytest.fixture(scope="module")
def solved(oracle_mod):
    o = make(oracle_mod)
    o.solve()
    return o


def test_dispersion_and_hoppings(oracle_mod):
    # hamiltonien.cpp:49-68,106-109,119-122
    o = make(oracle_mod, tp=-0.2, tpp=0.1, t3=0.05)
    nk = o.nk
    ek = o.get("ek").reshape(nk, nk)
    k = np.arange(nk) * PI / (nk - 1)
    kx, ky = np.meshgrid(k, k, indexing="ij")
    ref = (-2 * (np.cos(kx) + np.cos(ky)) - 4 * (-0.2) * np.cos(kx) * np.cos(ky)
           - 2 * 0.1 * (np.cos(2 * kx) + np.cos(2 * ky)) - 4 * 0.05 * (np.cos(2 * kx) * np.cos(ky) + np.cos(2 * ky) * np.cos(kx)))
    assert relerr(ek, ref) < 1e-15
    er = o.get("er")
    assert list(er[:5]) == [0.0, -1.0, 0.2, -0.1, -0.05]
    # eps(k) is the lattice Fourier transform of the real-space hoppings er on the wedge
    assert abs(ek.mean()) < 1.0  # sum over quadrant of cos terms is not zero; sanity only


@pytest.mark.parametrize("n,tp", [(0.875, -0.175), (1.0, -0.3), (1.1, -0.2), (0.6, 0.0)])
def test_mu0_reproduces_density(oracle_mod, n, tp):
    # green0.cpp:182-262: 2/N^2 sum_k f(eps_k - mu0) = n on the finite lattice
    o = make(oracle_mod, n=n, tp=tp)
    mu0 = o.s.mu0
    N = o.s.p.nk0
    k = 2 * PI * np.arange(N) / N
    kx, ky = np.meshgrid(k, k, indexing="ij")
    e = -2 * (np.cos(kx) + np.cos(ky)) - 4 * tp * np.cos(kx) * np.cos(ky)
    dens = 2 * np.mean(1.0 / (np.exp((e - mu0) / 0.1) + 1.0))
    assert abs(dens - n) < 1e-12


def test_mu0_half_filling_shortcut(oracle_mod):
    o = make(oracle_mod, n=1.0, tp=0.0)
    assert o.s.mu0 == 0.0


@pytest.mark.parametrize("Nw,tol", [(64, 1e-3), (256, 5e-6), (1024, 5e-8)])
def test_chi0_matches_brute_force_lindhard(oracle_mod, Nw, tol):
    # S1, chi0.cpp:744-753 with chiqw_discret (chi0.cpp:1588-1651); error falls as Nw^-4
    o = make(oracle_mod, Nw=Nw)
    o.calc_chiqw_FFT_spline(16, Nw)
    chi = o.get("chiqw").reshape(-1, o.nk8)
    for (l, m) in [(8, 8), (4, 2), (8, 0), (3, 3), (5, 1)]:
        idx = m + l * (l + 1) // 2
        for n in (0, 1, 3, 7):
            b = o.chiqw_discret(l, m, n)
            assert abs(chi[n, idx] - b) <= tol * abs(b)


def test_chi0_q0_is_zero_at_finite_frequency(oracle_mod):
    o = make(oracle_mod, Nw=128)
    o.calc_chiqw_FFT_spline(16, 128)
    chi = o.get("chiqw").reshape(-1, o.nk8)
    assert np.all(chi[1:, 0] == 0.0)
    assert chi[0, 0] > 0


def test_chi0_trace_sum_rule(oracle_mod):
    # S2, chi0.cpp:1498-1584: T/N^2 sum_{q,n} chi0 = n - n^2/2 (with tail correction)
    o = make(oracle_mod, Nw=1024)
    o.calc_chiqw_FFT_spline(16, 1024)
    tr, corr = o.trace_chi0()
    assert abs(corr - (0.875 - 0.875 ** 2 / 2)) < 5e-6
    assert abs(tr - (0.875 - 0.875 ** 2 / 2)) > abs(corr - (0.875 - 0.875 ** 2 / 2))


def test_usp_uch_satisfy_local_sum_rules(solved):
    # S3, chi.cpp:833-843: the roots satisfy the spin and charge sum rules
    s = solved.scalars()
    assert abs(solved.sum_rule(0, s["Usp"])) < 1e-14
    assert abs(solved.sum_rule(1, s["Uch"])) < 1e-14
    assert 0 < s["Usp"] < 4.0 < s["Uch"]  # Kanamori-Brueckner screening, charge enhancement
    assert s["Usp"] < 2.0 / s["chi0max"]
    assert abs(s["dblOcc"] - s["Usp"] / 4.0 * 0.875 ** 2 / 4.0) < 1e-15


def test_vertex_local_equal_time_value(solved):
    # S4, chi.cpp:177-179: V(r=0, tau=0) from the sum rules
    s = solved.scalars()
    U, nsr, Usp, Uch = 4.0, 0.875, s["Usp"], s["Uch"]
    exact = U * (3 * Usp * (nsr - Usp * nsr * nsr / (2 * U)) + Uch * (nsr + Usp * nsr * nsr / (2 * U) - nsr * nsr)) / 8
    assert abs(solved.get("V")[0] - exact) < 5e-5 * abs(exact)


def test_self_energy_matches_brute_force(solved):
    # S7, green.cpp:2034-2052 with Self_kw (green.cpp:1401-1494)
    s = solved.scalars()
    nk8 = solved.nk8
    Self = solved.get("Self").reshape(-1, nk8)
    T = 0.1
    for (l, m, n) in [(4, 4, 0), (8, 3, 1), (2, 0, 0), (6, 6, 2)]:
        idx = m + l * (l + 1) // 2
        z = 1j * (2 * n + 1) * PI * T
        S = s["Sigma_inf"] / z + solved.get("Sigma_inf2")[idx] / z ** 2 + solved.get("Sigma_inf3")[idx] / z ** 3 + Self[n, idx] / z ** 4
        B = solved.self_kw_brute(l, m, n)
        assert abs(S - B) < 2e-5 * abs(B)


def test_trace_sigma_G0_equals_potential_energy(solved):
    # S5, green.cpp:3086-3089: Tr(Sigma G0) = U <n_up n_dn> - U n^2/4
    s = solved.scalars()
    target = 4.0 * s["dblOcc"] - 4.0 * 0.875 ** 2 / 4
    assert abs(s["trSG_corr"] - target) < 2e-4 * abs(target)


def test_mu_root_and_interacting_density(solved, oracle_mod):
    # green.cpp:3108-3282: Tr(G - G0) = 0 at mu, summed up to n = Nw.  The exact density
    # of G then differs from n by the (mu0 - mu)/(i w_n)^2 tail beyond n = Nw that the
    # reference's trace leaves out: -(mu0 - mu) / (pi^2 T (Nw + 1)) (both spins).
    s = solved.scalars()
    assert abs(solved.traceDiffG(s["mu"])) < 1e-13
    dens, _ = _density_and_kx0(solved)
    T, Nw = solved.s.p.T, solved.s.nbw
    expected = 0.875 - (s["mu0"] - s["mu"]) / (PI ** 2 * T * (Nw + 1))
    assert abs(dens - expected) < 2e-5 * abs(s["mu0"] - s["mu"]) / (PI ** 2 * T * (Nw + 1)) + 2e-6


def _sigma_full(o, idx, n):
    """Sigma(k, i w_n) from the moment form, any integer n (negative: conjugate)."""
    s = o.scalars()
    nk8 = o.nk8
    Nw = o.s.nbw
    T = o.s.p.T
    nn = n if n >= 0 else -n - 1
    wn = (2 * nn + 1) * PI * T
    z = 1j * wn
    if nn < Nw // 2:
        R = o.get("Self").reshape(-1, nk8)[nn, idx]
    else:
        R = o.get("Sigma_inf4")[idx]
    v = s["Sigma_inf"] / z + o.get("Sigma_inf2")[idx] / z ** 2 + o.get("Sigma_inf3")[idx] / z ** 3 + R / z ** 4
    return v if n >= 0 else np.conj(v)


def _G(o, idx, e, nvals, two_pole_beyond=False):
    """G(k, i w_n) = 1/(i w_n - e + mu - Sigma); with two_pole_beyond, |n| >= Nw/2 uses
    Sigma = Sigma_inf/(i w_n) only (the asymptote of cond_opt.cpp:3368,3410-3457)."""
    s = o.scalars()
    mu, T, Nw = s["mu"], o.s.p.T, o.s.nbw
    nvals = np.asarray(nvals)
    wn = (2 * nvals + 1) * PI * T
    sig = np.array([_sigma_full(o, idx, n) for n in nvals])
    if two_pole_beyond:
        far = (nvals >= Nw // 2) | (nvals < -Nw // 2)
        sig[far] = s["Sigma_inf"] / (1j * wn[far])
    return 1.0 / (1j * wn - e + mu - sig)


def _density_and_kx0(o, two_pole_beyond=False):
    """n = 2/N^2 sum_k n(k), kx0 = -2/N^2 sum_k d2eps/dkx2 n(k); n(k) by direct Matsubara
    sums with G = 1/z + xi/z^2 + (xi^2 + S1)/z^3 + (C2 + 2 xi S1 + xi^3)/z^4 + O(z^-5)
    summed analytically: T sum_n e^{i w_n 0+}/z = 1/2, T sum 1/z^2 = -1/(4T), T sum 1/z^3
    = 0, T sum 1/z^4 = 1/(48 T^3) (z = i w_n, xi = eps - mu, S1 = Sigma_inf)."""
    N, T = o.s.p.nk0, o.s.p.T
    nk = o.nk
    Nw = o.s.nbw
    tp, tpp, t3 = o.s.p.tp, o.s.p.tpp, o.s.p.t3
    ek = o.get("ek").reshape(nk, nk)
    nvals = np.arange(-Nw, Nw)
    wn = (2 * nvals + 1) * PI * T
    dens = 0.0
    kx0 = 0.0
    for l in range(nk):
        for m in range(nk):
            x, y = max(l, m), min(l, m)
            idx = y + x * (x + 1) // 2
            G = _G(o, idx, ek[l, m], nvals, two_pole_beyond)
            xi = ek[l, m] - o.scalars()["mu"]
            z = 1j * wn
            S1 = o.scalars()["Sigma_inf"]
            c3 = xi ** 2 + S1
            c4 = (0.0 if two_pole_beyond else o.get("Sigma_inf2")[idx]) + 2 * xi * S1 + xi ** 3
            nkv = (T * np.sum(G - 1.0 / z - xi / z ** 2 - c3 / z ** 3 - c4 / z ** 4).real
                   + 0.5 - xi / (4 * T) + c4 / (48 * T ** 3))
            wx = 1 if l in (0, nk - 1) else 2
            wy = 1 if m in (0, nk - 1) else 2
            kxv, kyv = l * PI / (nk - 1), m * PI / (nk - 1)
            d2 = (2 * np.cos(kxv) + 4 * tp * np.cos(kxv) * np.cos(kyv) + 8 * tpp * np.cos(2 * kxv)
                  + 4 * t3 * (4 * np.cos(2 * kxv) * np.cos(kyv) + np.cos(2 * kyv) * np.cos(kxv)))
            dens += 2 * wx * wy * nkv / N ** 2
            kx0 += -2 * wx * wy * d2 * nkv / N ** 2
    return dens, kx0


def test_kx0_matches_direct_sum(solved):
    # cond_opt.cpp:3783 and 3993-4173: diamagnetic term from the occupation numbers, with
    # the reference's two-pole G beyond |n| = Nw/2
    _, kx0 = _density_and_kx0(solved, two_pole_beyond=True)
    assert abs(solved.scalars()["kx0"] - kx0) < 5e-7 * abs(kx0)


def test_bubble_matches_direct_matsubara_sum(solved):
    """Lambda_jj1(iq_m) = -2/N^2 sum_k wt v_x^2 T sum_n G(k,iw_n) G(k,iw_n+iq_m) -- the
    frequency-space form of the tau-space bubble (cond_opt.cpp:3734-3817), with the
    1/(i w_n)(i w_n + i q_m) tail summed analytically (-1/(4T) at m=0, 0 otherwise)."""
    o = solved
    N, T, nk, Nw = o.s.p.nk0, o.s.p.T, o.nk, o.s.nbw
    tp = o.s.p.tp
    ek = o.get("ek").reshape(nk, nk)
    c1 = o.get("chijj1")
    nvals = np.arange(-2 * Nw, 2 * Nw)
    wn = (2 * nvals + 1) * PI * T
    for mq in (0, 1, 2, 5):
        tot = 0.0
        for l in range(1, nk - 1):
            for m in range(nk):
                x, y = max(l, m), min(l, m)
                idx = y + x * (x + 1) // 2
                G = _G(o, idx, ek[l, m], np.arange(-2 * Nw, 2 * Nw + mq))
                G1, G2 = G[: 4 * Nw], G[mq: mq + 4 * Nw]
                a1, a2 = 1.0 / (1j * wn), 1.0 / (1j * (wn + 2 * mq * PI * T))
                s = T * np.sum(G1 * G2 - a1 * a2).real + (-1.0 / (4 * T) if mq == 0 else 0.0)
                kxv, kyv = l * PI / (nk - 1), m * PI / (nk - 1)
                v = 2 * np.sin(kxv) + 4 * tp * np.sin(kxv) * np.cos(kyv)
                wt = 4 if m not in (0, nk - 1) else 2
                tot += -2 * wt * v * v * s / N ** 2
        assert abs(c1[mq] - tot) < 2e-4 * abs(c1[0])


def test_fsum_and_positivity(solved):
    # S8 (cond_opt.cpp:4282-4286): Lambda1(0)+Lambda2(0) ~ -kx0 (approximate: TPSC is not
    # conserving); S9 (:4206-4214): the bubble is non-negative
    s = solved.scalars()
    c1, c2 = solved.get("chijj1"), solved.get("chijj2")
    assert abs((c1[0] + c2[0] + s["kx0"]) / s["kx0"]) < 0.05
    assert np.all(c1 >= 0)
    assert np.all(np.diff(c1[:20]) < 0)  # decreasing with q_n at low frequency


def test_particle_hole_symmetry(oracle_mod):
    o = make(oracle_mod, N=16, Nw=128, n=1.0, T=0.2, U=2.0, tp=0.0)
    o.solve()
    s = o.scalars()
    assert s["mu0"] == 0.0 and abs(s["mu"]) < 1e-11
    assert s["qmax"] == (8, 8)  # (pi, pi) nesting
    c1, c2 = o.get("chijj1"), o.get("chijj2")
    assert abs((c1[0] + c2[0] + s["kx0"]) / s["kx0"]) < 0.05


def test_density_above_half_filling_branch(oracle_mod):
    # chi.cpp:18-19,370-371: n > 1 uses nsr = 2 - n and adds 1 - nsr to the double occupancy
    o = make(oracle_mod, n=1.1, tp=-0.2)
    o.find_Usp_Uch()
    s = o.scalars()
    assert abs(s["nsr"] - 0.9) < 1e-15
    assert abs(s["dblOcc"] - (s["Usp"] / 4.0 * 0.81 / 4.0 + 1.0 - 0.9)) < 1e-15
    assert abs(o.sum_rule(0, s["Usp"])) < 1e-14


def test_coarse_pass_only_changes_diagnostics(oracle_mod):
    # chi.cpp:285-310: the N/2 x Nw/2 pass feeds difRelsp only
    a = make(oracle_mod, coarse=True)
    b = make(oracle_mod, coarse=False)
    a.find_Usp_Uch()
    b.find_Usp_Uch()
    sa, sb = a.scalars(), b.scalars()
    assert sa["Usp"] == sb["Usp"] and sa["Uch"] == sb["Uch"] and sa["mu0"] == sb["mu0"]
    assert sa["difRelsp"] > 0 and sa["Usp_coarse"] != sa["Usp"]


def test_hopping_t_is_an_energy_scale(oracle_mod):
    """eps(k; t, t') = t eps(k; 1, t'/t), and the whole chain is homogeneous in its energies:
    Sigma at (t, t', U, T, mu0) is t x Sigma at (1, t'/t, U/t, T/t, mu0/t) on the same Matsubara
    index, so L = sum_k w_k Re Sigma(k, i w_0) scales the same way.  Pins the generalization of
    eps(k), eps(r) (er[1] = -t) and the derivatives to a free hopping t."""
    import numpy as np

    N, Nw, n, T, U, tp, mu0, t = 16, 128, 0.9, 0.15, 2.5, -0.2, -0.3, 1.3
    w = np.linspace(0.2, 1.0, (N // 2 + 1) * (N // 2 + 2) // 2)
    a = oracle_mod.Oracle(N, Nw, n, T, U, tp=tp, coarse_pass=False, t=t, mu0=mu0).sigma_loss(w)
    b = oracle_mod.Oracle(N, Nw, n, T / t, U / t, tp=tp / t, coarse_pass=False, t=1.0, mu0=mu0 / t).sigma_loss(w)
    assert abs(a - t * b) <= 1e-9 * abs(a)
    # t = 0 in the parameters means the reference's t = 1
    c = oracle_mod.Oracle(N, Nw, n, T, U, tp=tp, coarse_pass=False, t=0.0, mu0=mu0).sigma_loss(w)
    d = oracle_mod.Oracle(N, Nw, n, T, U, tp=tp, coarse_pass=False, t=1.0, mu0=mu0).sigma_loss(w)
    assert c == d
