"""Lambda_jj by angle of k (N_theta_k > 1, cond_opt.cpp:3140-3174, 3792-3847, 6261-6398): the
bubble, Maki-Thompson and Aslamazov-Larkin contributions summed per angle bin of k, against the
oracle's binned arrays at 1e-10 (max |difference| over max |total|), and the bins adding up to the
totals."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

CASES = [  # (N, Nw, n, T, U, t', N_theta, mmin)
    (32, 256, 0.9, 0.1, 4.0, -0.175, 4, 2),
    (48, 128, 1.1, 0.15, 5.0, -0.24, 3, 1),
]


@pytest.mark.parametrize("case", CASES)
def test_binned_lambda_matches_oracle(case, oracle_mod):
    from tpsc_conductivity_amd.tpsc import CondOpt

    N, Nw, n, T, U, tp, nth, mmin = case
    print(f"[theta] N={N} Nw={Nw} n={n} T={T} U={U} tp={tp} N_theta={nth} mmin={mmin}", flush=True)
    o = oracle_mod.Oracle(N, Nw, n, T, U, tp=tp, coarse_pass=False)
    o.solve()
    o.calc_chijj(nth)
    o.calc_chijj_al(mmin, nth)
    g = CondOpt(N, Nw, n, T, params=(tp, 0.0, 0.0, U), coarse_pass=False)
    g.solve()
    g.calc_chijj_vertex_corr_all_optim(chi3=True, mmin=mmin, N_theta_k=nth)
    assert g.scalar("n_theta") == nth
    nwb = Nw // 2 + 1
    for f in ("chijj1_k", "chijj2_k"):
        a, b = g.get(f), o.get(f)
        assert a.size == nwb * nth
        e = float(np.abs(a - b).max() / np.abs(o.get(f[:-2])).max())
        print(f"[theta]   {f} rel err {e:.2e}", flush=True)
        assert e < 1e-10
        tot = a.reshape(nwb, nth).sum(axis=1)
        assert relerr(tot, g.get(f[:-2])) < 1e-13
    a3, b3 = g.get("chijj3_k"), o.get("chijj3_k")
    # normwise against the totals: each bin carries the absolute rounding error of a partial sum
    e = float(np.abs(a3 - b3).max() / np.abs(o.get("chijj3")).max())
    print(f"[theta]   chijj3_k rel err {e:.2e}", flush=True)
    assert e < 1e-10
    assert relerr(a3.reshape(-1, nth).sum(axis=1), g.get("chijj3")) < 1e-12
    # every bin holds some k points, so no bin is trivially zero
    assert np.all(np.abs(a3.reshape(-1, nth)[1:]).max(axis=0) > 0)
    g.close()


def test_single_bin_equals_the_totals():
    from tpsc_conductivity_amd.tpsc import CondOpt

    g = CondOpt(16, 64, 0.9, 0.2, params=(-0.1, 0.0, 0.0, 3.0), coarse_pass=False)
    g.solve()
    assert np.array_equal(g.get("chijj1_k"), g.get("chijj1"))
    assert np.array_equal(g.get("chijj2_k"), g.get("chijj2"))
    g.close()
