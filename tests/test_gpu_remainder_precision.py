"""The raw 1/z^4 remainder R(k, n) of Sigma(k, i w_n), which the parity tests compare with the oracle
at 1e-8 instead of 1e-10, recomputed in x87 extended precision (tests/remainder_ld.py) from the GPU's
own inputs (Sigma(k, tau_j) for 0 < j < Nw, the end samples and slopes the spline uses): the GPU's R
is that transform of its inputs to ~1e-12, as the oracle's R is of the oracle's inputs
(tests/test_oracle_numerics.py).  The two sides' inputs differ in their last bits only; the
transform's third-derivative sums amplify those differences at high n, which is the 1e-8 gap."""
import numpy as np
import pytest

import remainder_ld as rl
from conftest import relerr

pytestmark = pytest.mark.gpu

CASES = [(16, 256, 0.1), (32, 1024, 0.05)]


@pytest.mark.parametrize("N,Nw,T", CASES)
def test_remainder_is_the_exact_transform_of_the_gpu_inputs(N, Nw, T, oracle_mod):
    from tpsc_conductivity_amd.tpsc import CondOpt

    print(f"[remainder] N={N} Nw={Nw} T={T}", flush=True)
    g = CondOpt(N, Nw, 0.875, T, params=(-0.175, 0.0, 0.0, 4.0), coarse_pass=False)
    g.calc_Self_FFT_spline()
    nk8 = (N // 2 + 1) * (N // 2 + 2) // 2
    St = g.get("Sigma_tau").reshape(Nw + 1, nk8).copy()
    Sb = g.get("Sigma_bnd").reshape(2, nk8)
    St[0], St[Nw] = Sb[0], Sb[1]  # the end samples the spline uses
    dS = g.get("dtau_Sigma").reshape(2, nk8)
    R = g.get("Self").reshape(-1, nk8)[: Nw // 2]
    ks = np.arange(nk8) if nk8 < 60 else np.linspace(0, nk8 - 1, 8).astype(int)
    Rl = rl.remainder(St[:, ks], dS[0, ks], dS[1, ks], T).astype(np.complex128)
    e = np.max(np.abs(R[:, ks] - Rl), axis=0) / np.max(np.abs(Rl), axis=0)
    o = oracle_mod.Oracle(N, Nw, 0.875, T, 4.0, tp=-0.175, coarse_pass=False)
    assert o.solve()
    Ro = o.get("Self").reshape(-1, nk8)[: Nw // 2]
    gap_in = relerr(g.get("Sigma_tau"), o.get("Sigma_tau"))
    gap_R = np.max(np.abs(R[:, ks] - Ro[:, ks]), axis=0) / np.max(np.abs(Ro[:, ks]), axis=0)
    print(f"[remainder] GPU R vs its extended-precision transform: {e.max():.1e}; inputs GPU vs oracle "
          f"{gap_in:.1e}; R GPU vs oracle {gap_R.max():.1e}", flush=True)
    assert e.max() < 1e-11
    assert gap_in < 1e-13
    g.close()
