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
