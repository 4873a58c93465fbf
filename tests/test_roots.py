"""Host root search of the product library (tpsc_find_zero, no device needed).

The library's bracketed search (Chandrupatla's inverse-quadratic / bisection hybrid with a
bracket-widening step) replaces the reference's find_zero (generique.cpp:604-913, :917-1234).
Pinned by closed-form roots, by agreement with the oracle's restatement of the reference's
search on the functions the path solves (finite-lattice density, spin sum rule, Tr(G - G0)),
and by its failure semantics (no sign change inside the limits, a flat function).
"""
import math

import numpy as np

from tpsc_conductivity_amd import capi


def test_closed_form_roots():
    cases = [
        (lambda x: x ** 3 - 2 * x - 5, [1.0, 0.0, 4.0], 2.0945514815423265),
        (lambda x: math.cos(x) - x, [0.5, 0.0, 1.0], 0.7390851332151607),
        (lambda x: math.exp(x) - 10.0, [0.0, -5.0, 5.0], math.log(10.0)),
        # a pole at the top of the bracket, like the spin sum rule near 2/chi0max
        (lambda x: 1.0 / (3.0 - x) - 2.0, [1.0, 0.0, 3.0 - 1e-12], 2.5),
        # triple root: derivative-free bracketing still converges
        (lambda x: (x - 1.25) ** 3, [0.0, -1.0, 4.0], 1.25),
    ]
    for f, init, ref in cases:
        ok, r, ne = capi.find_zero(f, init)
        assert ok, (init, r)
        assert abs(r - ref) < 1e-13 * max(1.0, abs(ref)), (init, r, ref)
        assert ne < 120
        ok, r, _ = capi.find_zero(f, init, lims=[-20.0, 20.0])
        assert ok and abs(r - ref) < 1e-13 * max(1.0, abs(ref))


def test_widening_and_limits():
    g = lambda x: (x - 5.5) * (1.0 + 0.1 * x * x)
    # the root lies outside [1, 3]: the bracket grows until it changes sign
    ok, r, _ = capi.find_zero(g, [0.0, 1.0, 3.0])
    assert ok and abs(r - 5.5) < 1e-13
    # with limits the bracket may not leave them: no root inside [-5, 5]
    ok, _, _ = capi.find_zero(g, [0.0, 1.0, 3.0], lims=[-5.0, 5.0])
    assert not ok
    # inside the limits it is found
    ok, r, _ = capi.find_zero(g, [0.0, 1.0, 3.0], lims=[-20.0, 20.0])
    assert ok and abs(r - 5.5) < 1e-13
    # a function without a root is reported as a failure
    ok, _, _ = capi.find_zero(lambda x: 1.0 + x * x, [0.0, -1.0, 1.0])
    assert not ok
    # a flat function whose sign change lies beyond the widening reach fails too
    ok, _, _ = capi.find_zero(lambda x: np.tanh(x - 1e30), [0.0, 1.0, 3.0])
    assert not ok


def test_two_roots_keeps_upper_side(oracle_mod):
    # f(lo), f(hi) > 0 and f(start) < 0: both halves bracket a root; the search keeps
    # [start, hi], the side the reference's search takes (generique.cpp:627-633)
    f = lambda x: (x - 1.0) * (x - 3.0)
    ok, r, _ = capi.find_zero(f, [2.0, 0.0, 4.0])
    assert ok and abs(r - 3.0) < 1e-14
    ok2, r2, _ = oracle_mod.find_zero(f, [2.0, 0.0, 4.0])
    assert ok2 and abs(r - r2) < 1e-14


def test_same_roots_as_reference_search(oracle_mod):
    """The functions the TPSC path solves: the library's search and the oracle's restatement
    of the reference's search agree to rounding."""
    o = oracle_mod.Oracle(16, 64, 0.875, 0.1, 4.0, tp=-0.175, coarse_pass=False)
    # finite-lattice density (green0.cpp:182-203, start {0, -8, 8})
    f = lambda m: oracle_mod.lib().orc_ddensity_mu_PBC(o._o, m)
    ok, r, _ = capi.find_zero(f, [0.0, -8.0, 8.0])
    ok2, r2, _ = oracle_mod.find_zero(f, [0.0, -8.0, 8.0])
    assert ok and ok2 and abs(r - r2) < 1e-13
    # spin sum rule on (0, 2/chi0max), start U/2 (chi.cpp:853-858)
    o.find_mu0()
    o.calc_chiqw_FFT_spline(16, 64)
    c0max = o.scalars()["chi0max"]
    g = lambda u: o.sum_rule(0, u)
    init = [2.0, 0.0, 2.0 / c0max - 2 * np.finfo(float).eps]
    ok, r, _ = capi.find_zero(g, init)
    ok2, r2, _ = oracle_mod.find_zero(g, init)
    assert ok and ok2 and abs(r - r2) < 1e-12 * abs(r2)
    # charge sum rule at that Usp on (0, 1000 U), start 0 (chi.cpp:862-869)
    o.find_Usp_Uch()
    h = lambda u: o.sum_rule(1, u)
    init = [0.0, 0.0, 4000.0]
    ok, r, _ = capi.find_zero(h, init)
    ok2, r2, _ = oracle_mod.find_zero(h, init)
    assert ok and ok2 and abs(r - r2) < 1e-12 * abs(r2)
    # Tr(G - G0)(mu) with limits +-20 (green.cpp:3238)
    o.calc_chirt()
    o.calc_Self()
    d = lambda m: o.traceDiffG(m)
    ok, r, _ = capi.find_zero(d, [0.0, -10.0, 10.0], lims=[-20.0, 20.0])
    ok2, r2, _ = oracle_mod.find_zero(d, [0.0, -10.0, 10.0], lims=[-20.0, 20.0])
    assert ok and ok2 and abs(r - r2) < 1e-12 * max(1.0, abs(r2))


# batch point 41 of tpsc_batch_points(64064, 64): 1 - Usp chi0max/2 ~ 3e-5, the spin sum rule
# rises by ~5.7e3 per unit of U just below its pole at 2/chi0max, the charge sum rule falls by
# only 1.6e-3 per unit of U at Uch ~ 45
NEAR_POLE = dict(N=128, Nw=128, n=0.9690175914152073, T=0.0672362960301115, U=4.676953119127109,
                 tp=-0.040868229371670894)


def newton_step(f, x):
    h = 1e-6 * abs(x)
    return f(x) / ((f(x + h) - f(x - h)) / (2 * h))


def test_reference_search_stops_short_near_the_instability(oracle_mod):
    """At NEAR_POLE the reference's search (generique.cpp:604-913: 20 iterations, then the point
    of smallest |f| kept) ends 2.3e-7 below the Usp root (|f| ~ 1.8e-3), and, at the converged
    Usp, one Newton step (2.4e-10 relative) short of the Uch root; the library's search
    converges on both.  Device results there are checked against the oracle's chain rerun from
    the converged roots (orc_solve_from_roots)."""
    p = NEAR_POLE
    o = oracle_mod.Oracle(p["N"], p["Nw"], p["n"], p["T"], p["U"], tp=p["tp"], coarse_pass=False)
    assert o.solve()
    s = o.scalars()
    f = lambda u: o.sum_rule(0, u)
    assert s["nevals_sp"] == 40 and abs(f(s["Usp"])) > 1e-3
    init = [p["U"] / 2, 0.0, 2.0 / s["chi0max"] - 2 * np.finfo(float).eps]
    ok, usp, _ = capi.find_zero(f, init)
    assert ok and abs(newton_step(f, usp)) < 1e-14 * usp and 2e-7 < (usp - s["Usp"]) / usp < 3e-7
    assert o.solve_from_roots(usp)  # Uch searched the reference's way at the converged Usp
    g = lambda u: o.sum_rule(1, u)
    uch_ref = o.scalars()["Uch"]
    assert 1e-10 < abs(newton_step(g, uch_ref)) / uch_ref < 1e-9
    ok, uch, _ = capi.find_zero(g, [0.0, 0.0, 1000.0 * p["U"]])
    assert ok and abs(newton_step(g, uch)) < 1e-13 * uch


def test_chain_from_given_roots_reproduces_the_solve(oracle_mod):
    o = oracle_mod.Oracle(32, 128, 0.9, 0.1, 4.0, tp=-0.175, coarse_pass=False)
    assert o.solve()
    s0, j0 = o.scalars(), o.get("chijj1") + o.get("chijj2")
    for args in ((s0["Usp"],), (s0["Usp"], s0["Uch"])):
        assert o.solve_from_roots(*args)
        s1, j1 = o.scalars(), o.get("chijj1") + o.get("chijj2")
        for k in ("Usp", "Uch", "dblOcc", "mu", "kx0"):
            assert s1[k] == s0[k], (k, args)
        assert np.array_equal(j0, j1)
    # a different Usp moves every later stage, Uch included when it is searched
    assert o.solve_from_roots(s0["Usp"] * (1 + 1e-6))
    assert o.scalars()["Uch"] != s0["Uch"] and o.scalars()["mu"] != s0["mu"]
    assert o.solve_from_roots(s0["Usp"], s0["Uch"] * (1 + 1e-6))
    assert o.scalars()["Uch"] == s0["Uch"] * (1 + 1e-6) and o.scalars()["mu"] != s0["mu"]
