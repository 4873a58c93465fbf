"""Batches of independent parameter points (tpsc_solve_batch): every point's result equals the
same point solved alone (the solve is deterministic), unsupported lattices are reported per
point, and the worker count the tuner picks is one of its candidates."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def params(N, Nw, n, T, U, tp):
    from tpsc_conductivity_amd.capi import Params

    return Params(N, Nw, n, T, 0.1, tp, 0.0, 0.0, U, 0, 0, 1.0)


POINTS = [(32, 128, 0.9, 0.15, 3.0, -0.2), (16, 64, 0.875, 0.1, 4.0, -0.175), (8, 32, 0.95, 0.25, 2.0, 0.0),
          (32, 64, 0.85, 0.3, 5.0, -0.1), (16, 256, 1.0, 0.05, 2.5, 0.0), (24, 64, 0.9, 0.2, 2.0, 0.0),
          (20, 64, 0.9, 0.2, 2.0, 0.0)]


def test_batch_matches_single_solves():
    from tpsc_conductivity_amd import capi
    from tpsc_conductivity_amd.tpsc import CondOpt

    pts = (capi.Params * len(POINTS))(*[params(*p) for p in POINTS])
    print("[batch] 7 points, 3 workers", flush=True)
    res, jj = capi.solve_batch(pts, workers=3, ldc=129)
    assert res[6].status == -1          # N = 20: neither 2^k nor 3*2^k
    for i, (N, Nw, n, T, U, tp) in enumerate(POINTS[:6]):
        assert res[i].status == 0, i
        g = CondOpt(N, Nw, n, T, params=(tp, 0.0, 0.0, U), coarse_pass=False)
        g.solve()
        assert res[i].Usp == g.scalar("Usp") and res[i].mu == g.scalar("mu") and res[i].kx0 == g.scalar("kx0")
        ref = g.get("chijj1") + g.get("chijj2")
        assert np.array_equal(jj[i, : Nw // 2 + 1], ref)
        g.close()


def test_batch_tuner_and_the_random_point_set():
    from tpsc_conductivity_amd import capi

    pts = capi.batch_points(64064, 64)
    assert all(p.Nw0 >= 16 and (p.Nw0 - 1) * np.pi * p.T >= 20 for p in pts)
    w, rates = capi.batch_tune(pts)
    assert w in (1, 2, 4, 8) and np.all(rates > 0)
    print(f"[batch] tuner: workers {w}, points/s {rates}", flush=True)
    res, _ = capi.solve_batch(pts, workers=w)
    st = np.array([r.status for r in res])
    assert {p.nk0 for p in pts} == {32, 48, 64, 96, 128}
    assert np.all(st == 0)
