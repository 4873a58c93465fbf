"""Host post-processing of a solved point through the C ABI (tpsc_mu0_infinite, tpsc_fermi_surface,
tpsc_fs_lattice, tpsc_mdc, tpsc_edc, tpsc_self_real_freq, the binary restart files) against the
oracle's restatement of the reference (oracle/post.py), fed the same device results (eps(k), mu,
Sigma(k, iw_n) from tpsc_get).  oracle/post.py itself is pinned in tests/test_oracle_post.py."""
import os

import numpy as np
import pytest

import post
from conftest import relerr

pytestmark = pytest.mark.gpu

CFG = dict(N=32, Nw=256, n=0.9, T=0.1, U=4.0, tp=-0.175, tpp=0.05)


@pytest.fixture(scope="module")
def solved():
    from tpsc_conductivity_amd.tpsc import CondOpt

    c = CFG
    print(f"[post] solve {c}", flush=True)
    g = CondOpt(c["N"], c["Nw"], c["n"], c["T"], params=(c["tp"], c["tpp"], 0.0, c["U"]), coarse_pass=False)
    g.solve()
    nk = c["N"] // 2 + 1
    nk8 = nk * (nk + 1) // 2
    H = c["Nw"] // 2
    d = dict(g=g, nk=nk, nk8=nk8, H=H, ek=g.get("ek"), mu=g.scalar("mu"), mu0=g.scalar("mu0"),
             S=g.get("Sigma_kw").reshape(H, nk8), R=g.get("Self").reshape(H, nk8))
    yield d
    g.close()


def test_infinite_lattice_mu0(solved):
    c = CFG
    got = solved["g"].find_mu_inf_syst()
    want = post.mu0_infinite(c["n"], c["T"], c["tp"], c["tpp"], 0.0)
    print(f"[post] mu0 infinite {got!r} oracle {want!r}, finite {solved['mu0']!r}", flush=True)
    assert abs(got - want) < 1e-11


def test_fermi_surfaces(solved):
    c, g = CFG, solved["g"]
    fs = g.find_FS(1001)
    ref = post.fermi_surface(solved["mu0"], c["tp"], c["tpp"], 0.0, nk=1001)
    assert fs.shape == ref.shape and len(fs) > 100
    assert np.abs(fs - ref).max() < 1e-12
    m, a = g.find_FS_PBC()
    mr, ar = post.fs_lattice(solved["ek"], solved["nk"], solved["mu0"])
    assert np.array_equal(m, mr)
    assert relerr(a, ar) < 1e-12
    (mp, ap), (mq, aq) = g.find_FS_inter()
    (rp, rap), (rq, raq) = post.fs_inter(solved["ek"], solved["nk"], solved["mu"], solved["S"], c["T"])
    assert np.array_equal(mp, rp) and np.array_equal(mq, rq)
    assert relerr(ap, rap) < 1e-10 and relerr(aq, raq) < 1e-10


@pytest.mark.parametrize("use_pade,NP", [(True, 10), (False, 4)])
def test_mdc(solved, use_pade, NP):
    g, S = solved["g"], solved["S"]
    etas = [0.01, 0.001]
    sw, ak = g.MDC(list(range(solved["H"])), NP, 0.05, etas, use_Pade=use_pade)
    rs, ra = post.mdc(solved["ek"], solved["nk"], solved["mu"], lambda j, k: S[j, k], np.arange(NP), CFG["T"], 0.05,
                      etas, use_pade=use_pade)
    e1, e2 = relerr(sw, rs), relerr(ak, ra)
    print(f"[post] MDC pade={use_pade}: Sigma {e1:.1e}, A {e2:.1e}", flush=True)
    assert e1 < 1e-9 and e2 < 1e-9


def test_edc(solved):
    g, S = solved["g"], solved["S"]
    kx, ky = [8, 12, -3, 20], [0, 5, 7, 30]
    w = np.linspace(-1.5, 1.5, 31)
    out = g.EDC(kx, ky, w, 1, 12, [0.05, 0.2])
    ref = post.edc(solved["ek"], solved["nk"], solved["mu"], lambda j, k: S[j, k], CFG["T"], CFG["Nw"], kx, ky, w, 1,
                   12, [0.05, 0.2])
    e = relerr(out, ref)
    print(f"[post] EDC {e:.1e}", flush=True)
    assert e < 1e-9


def test_self_real_freq(solved):
    g, S, R = solved["g"], solved["S"], solved["R"]
    k, n = (9, 4), list(range(16))
    w = np.linspace(-1.0, 1.0, 21)
    got = g.calc_Self_Re_w(k, n, w, [0.02, 0.1])
    ref = post.self_real_freq(solved["ek"], solved["nk"], solved["mu"], g.scalar("Sigma_inf"),
                              g.get("Sigma_inf2"), g.get("Sigma_inf3"), lambda j, kk: R[j, kk], CFG["T"], k, n, w,
                              [0.02, 0.1])
    for a, b in zip(got, ref):
        assert relerr(a, b) < 1e-9


def test_binary_restart_files(solved, tmp_path):
    from tpsc_conductivity_amd import capi

    c, g = CFG, solved["g"]
    g.save_Self_bin(str(tmp_path))
    g.save_chi(str(tmp_path))
    a8 = (c["U"], c["tp"], c["tpp"], 0.0, c["n"], c["T"], c["N"], c["Nw"])
    fself = tmp_path / ("self_bin_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8)
    fchi = tmp_path / ("chi0_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%d_%d.dat" % a8[1:])
    assert sorted(os.listdir(tmp_path)) == sorted([fself.name, fchi.name])
    nk8, H = solved["nk8"], solved["H"]
    assert os.path.getsize(fself) == (H + 1) * nk8 * 16
    arr = capi.load_Self_bin(fself, c["N"], c["Nw"])
    assert np.array_equal(arr[:H], solved["R"])
    st = g.get("Sigma_tau").reshape(c["Nw"] + 1, nk8)
    assert np.array_equal(arr[H].real, st[H]) and np.array_equal(arr[H].imag, st[c["Nw"]])
    assert os.path.getsize(fchi) == 8 + (H + 1) * nk8 * 8
    nbq, nwmax, chi = capi.loadChi0(fchi)
    assert (nbq, nwmax) == (solved["nk"], H)
    assert np.array_equal(chi.ravel(), g.get("chi0"))
