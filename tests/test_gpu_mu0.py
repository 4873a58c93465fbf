"""mu0 (green0::find_mu) is searched on the device and read by the host only when needed: a
fixed mu0 (tpsc_fix_mu0) must hold from the call on, also when the search queued by the
context's creation is still in flight, and releasing it (NaN) must give the search's root."""
import math

import pytest

pytestmark = pytest.mark.gpu

CFG = dict(N=32, Nw=128, n=0.9, T=0.1, U=3.0, tp=-0.2)


def make():
    from tpsc_conductivity_amd.tpsc import CondOpt

    c = CFG
    return CondOpt(c["N"], c["Nw"], c["n"], c["T"], params=(c["tp"], 0.0, 0.0, c["U"]), coarse_pass=False)


def test_fixed_mu0_holds_over_the_queued_search():
    g0 = make()
    root = g0.scalar("mu0")
    g0.close()
    x = root + 0.01
    print(f"[mu0] root {root:.15f}, fixed {x:.15f}", flush=True)
    g = make()  # its mu0 search is queued, not read
    g.fix_mu0(x)
    assert g.scalar("mu0") == x
    g.solve()
    assert g.scalar("mu0") == x
    g.fix_mu0(math.nan)
    assert g.scalar("mu0") == root
    g.close()


def test_fix_mu0_drops_every_stage_and_release_matches_fresh_context():
    """fix_mu0 (fixed value or NaN) resets every computed stage; a solve after the release
    gives the fresh context's results bit for bit."""
    import numpy as np

    g0 = make()
    g0.solve()
    ref = {k: g0.scalar(k) for k in ("mu0", "Usp", "Uch", "mu", "kx0")}
    ref_jj = (g0.get("chijj1"), g0.get("chijj2"))
    g0.close()
    g = make()
    g.solve()
    assert g.scalar("stages") != 0
    g.fix_mu0(ref["mu0"] + 0.02)
    assert g.scalar("stages") == 0
    g.solve()
    assert g.scalar("stages") != 0
    assert g.scalar("Usp") != ref["Usp"]
    g.fix_mu0(math.nan)
    assert g.scalar("stages") == 0
    assert g.scalar("mu0") == ref["mu0"]
    g.solve()
    for k, v in ref.items():
        assert g.scalar(k) == v, k
    assert np.array_equal(g.get("chijj1"), ref_jj[0])
    assert np.array_equal(g.get("chijj2"), ref_jj[1])
    g.close()

