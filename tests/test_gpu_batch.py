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


SCALARS = ("mu0", "Usp", "Uch", "mu", "dblOcc", "kx0")


@pytest.mark.parametrize("workers", [1, 3])
def test_batch_matches_single_solves(workers):
    """workers = 1 takes the queued path (no host wait inside a point), 3 the threaded one."""
    from tpsc_conductivity_amd import capi
    from tpsc_conductivity_amd.tpsc import CondOpt

    pts = (capi.Params * len(POINTS))(*[params(*p) for p in POINTS])
    print(f"[batch] 7 points, {workers} workers", flush=True)
    res, jj = capi.solve_batch(pts, workers=workers, ldc=129)
    assert res[6].status == -1          # N = 20: neither 2^k nor 3*2^k
    queued = sum(r.queued for r in res)
    print(f"[batch] {queued} of 6 points finished on the queued path", flush=True)
    assert queued >= 3 if workers == 1 else queued == 0
    for i, (N, Nw, n, T, U, tp) in enumerate(POINTS[:6]):
        assert res[i].status == 0, i
        g = CondOpt(N, Nw, n, T, params=(tp, 0.0, 0.0, U), coarse_pass=False)
        g.solve()
        for k in SCALARS:
            assert getattr(res[i], k) == g.scalar(k), (i, k)
        ref = g.get("chijj1") + g.get("chijj2")
        assert np.array_equal(jj[i, : Nw // 2 + 1], ref)
        g.close()


@pytest.mark.parametrize("precision,cap", [(0, None), (1, None), (1, 3)])
def test_queued_batch_equals_threaded_batch(precision, cap, monkeypatch):
    """The 64 random points, FP64 and mixed: the queued path (workers = 1) gives the threaded
    path's results bit for bit.  Mixed points near the antiferromagnetic instability (enhancement
    above 8) are queued a second time with the FP64 chi0 stage, as tpsc_solve redoes them.  With
    the searches capped at 3 passes (TPSC_QUEUE_MAX_PASSES) most points leave the queue and are
    solved again with tpsc_solve: the fallback path, same results."""
    from tpsc_conductivity_amd import capi

    pts = capi.batch_points(64064, 64)
    for p in pts:
        p.precision = precision
    if cap:
        monkeypatch.setenv("TPSC_QUEUE_MAX_PASSES", str(cap))
    r1, j1 = capi.solve_batch(pts, workers=1, ldc=129)
    monkeypatch.delenv("TPSC_QUEUE_MAX_PASSES", raising=False)
    if not cap:  # the queued path's contexts persist between calls; released, they are made again
        capi.batch_release()
        r2, j2 = capi.solve_batch(pts, workers=1, ldc=129)
        r2b, j2b = capi.solve_batch(pts, workers=1, ldc=129)
        for r in (r2, r2b):
            assert all(getattr(a, k) == getattr(b, k) for a, b in zip(r1, r) for k in SCALARS)
        assert np.array_equal(j1, j2) and np.array_equal(j1, j2b)
    r3, j3 = capi.solve_batch(pts, workers=3, ldc=129)
    queued = sum(r.queued for r in r1)
    print(f"[batch] precision {precision}, passes capped at {cap}: {queued} of 64 points finished on the "
          "queued path", flush=True)
    assert queued < 32 if cap else queued >= 60
    assert not any(r.queued for r in r3)
    for i in range(64):
        assert r1[i].status == 0 and r3[i].status == 0, i
        for k in SCALARS:
            assert getattr(r1[i], k) == getattr(r3[i], k), (i, k, r1[i].queued)
    assert np.array_equal(j1, j3)


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
