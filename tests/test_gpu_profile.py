"""Stage timing (tpsc_prof_enable): a solve with the library's HIP-event regions on the three
streams must run through, report every stage of the chain, and give the same results as a
solve without them (the regions only add events and waits)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG = dict(N=64, Nw=256, n=0.9, T=0.1, U=3.5, tp=-0.2, tpp=0.05, t3=0.0)
STAGES = ("chi0_tau0", "chi0_slab", "chi0_pencil", "sum_rules", "sigma_slab", "sigma_pencil", "mu_root",
          "jj_hf0_slab", "jj_hVf0V_slab", "jj_pencil")


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_profiled_solve(prec):
    from tpsc_conductivity_amd.tpsc import CondOpt

    c = CFG
    print(f"[profile] {prec} {c['N']}x{c['Nw']}", flush=True)
    g = CondOpt(c["N"], c["Nw"], c["n"], c["T"], params=(c["tp"], c["tpp"], c["t3"], c["U"]), coarse_pass=False,
                precision=prec)
    g.solve()
    ref = {k: g.scalar(k) for k in ("mu", "Usp", "Uch", "kx0")}
    jj = g.get("chijj1") + g.get("chijj2")
    g.profile(True)
    g.profile_reset()
    g.solve()
    table = g.profile_table()
    g.profile(False)
    g.solve()  # timing off again on the same context
    for k, v in ref.items():
        assert g.scalar(k) == v, k
    assert np.array_equal(g.get("chijj1") + g.get("chijj2"), jj)
    g.close()
    missing = [s for s in STAGES if s not in table]
    assert not missing, (missing, sorted(table))
    assert all(ms >= 0.0 and calls >= 1 for ms, calls in table.values())
