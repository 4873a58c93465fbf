"""Pins of the oracle's local Green function G_loc(iw_n) (green.cpp:209-266).

* The packed-wedge weights (8, halved on the edges and the diagonal) reproduce the plain sum
  over all N x N lattice points of 1/(iw_n - e_k + mu - Sigma(k,iw_n)), with Sigma(k) taken
  from the wedge through the lattice symmetries.
* G_loc(iw) = 1/(iw) + O(1/w^2): iw G_loc(iw) -> 1 at the top of the Matsubara range.
"""
import numpy as np


def test_gloc_wedge_weights_and_tail(oracle_mod):
    N, Nw = 16, 128
    o = oracle_mod.Oracle(N, Nw, 0.875, 0.1, 4.0, tp=-0.175, coarse_pass=False)
    assert o.solve()
    s = o.s
    T, mu, H = s.p.T, s.mu, Nw // 2
    nk = N // 2 + 1
    nk8 = nk * (nk + 1) // 2
    got = o.Gwn_local()
    R = o.get("Self")[: H * nk8].reshape(H, nk8)
    S2, S3, Sinf = o.get("Sigma_inf2"), o.get("Sigma_inf3"), o.scalars()["Sigma_inf"]
    z = 1j * (2 * np.arange(H) + 1) * np.pi * T
    ks = 2 * np.pi * np.arange(N) / N
    ref = np.zeros(H, dtype=complex)
    for ix in range(N):
        for iy in range(N):
            x, y = min(ix, N - ix), min(iy, N - iy)
            if y > x:
                x, y = y, x
            k = y + x * (x + 1) // 2
            sig = Sinf / z + S2[k] / z ** 2 + S3[k] / z ** 3 + R[:, k] / z ** 4
            ref += 1.0 / (z - o.dispersion(ks[ix], ks[iy]) + mu - sig)
    ref /= N * N
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    tail = np.abs(z * got - 1.0)
    assert tail[-1] < 0.05 and tail[-1] < tail[H // 4]
