"""The input.dat driver and tpsc_write_files on the GPU: the reference's file set, names and
formats (hamiltonien.cpp:72-82, green0.cpp:51-105, 626-676, chi0.cpp:720-737, chi.cpp:375-415,
green.cpp:942-1065, 2210-2341, 2978-2994, 3088-3103, 3259-3276, cond_opt.cpp:3889-3946, 4195-4279,
6447-6481), with the same numbers the C ABI returns for the same parameter point."""
import os
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "tpsc_conductivity_amd", "tpsc_run")
U, tp, tpp, t3, n, T, N, Nw, mmin = 4.0, -0.175, 0.0, 0.0, 0.875, 0.1, 16, 128, 2


def names(qx, qy):
    a8 = (U, tp, tpp, t3, n, T, N, Nw)
    return {
        "ek": "e_k_%5.3f_%5.3f_%5.3f_%1d.dat" % (tp, tpp, t3, N),
        "chi0w": "chi0_qmax_wn_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%d_%d_%d_%d.dat" % (tp, tpp, t3, n, T, N, Nw, qx, qy),
        "chi0q": "chi0_wn0_q_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%d_%d.dat" % (tp, tpp, t3, n, T, N, Nw),
        "infos": "TPSC_infos_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%d_%d.dat" % a8,
        "params": "TPSCparams_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%d_%d.dat" % a8,
        "mu": "mu0_mu_D_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8,
        "gloc": "Gwn_loc_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%d_%d.dat" % a8,
        "kx0": "kx0_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8,
        "kx0b": "kx0_bin_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8,
        "jj1": "chijj1_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8,
        "jj2": "chijj2_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8,
        "jjv": "chijj_with_vc1_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8,
        "jj3": "chijj3_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d_%1d.dat" % (a8 + (mmin,)),
        "jja": "chijj_with_all_vc_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d_%1d.dat" % (a8 + (mmin,)),
        "fs": "FS_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d.dat" % (tp, tpp, t3, n, T, 10001),
        "fsdA": "FS_discret_AkF_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d.dat" % (tp, tpp, t3, n, T, N),
        "fsd": "FS_discret_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d.dat" % (tp, tpp, t3, n, T, N),
        "fsiA": "FS_inter_AkF_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8,
        "fsi": "FS_inter_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8,
        "fsiAP": "FS_inter_AkF_Pade_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8,
        "fsiP": "FS_inter_Pade_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8,
        "mdcS": "Self_Pade_k_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%5.3e_%1d_%1d.dat" % (a8[:6] + (0.0, N, Nw)),
        "mdcA": "Akw_Pade_k_%4.2f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%5.3e_%1d_%1d.dat" % (a8[:6] + (0.0, N, Nw)),
        "jj1k": "chijj1_k_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8,
        "jj1kt": "chijj1_k_tot_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8,
        "jj2k": "chijj2_k_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8,
        "jj2kt": "chijj2_k_tot_qn_%5.3f_%5.3f_%5.3f_%5.3f_%7.5f_%6.4f_%1d_%1d.dat" % a8,
    }


def table(path, cols):
    rows = [ln.split() for ln in open(path) if ln.strip()]
    assert all(len(r) == cols for r in rows)
    return rows


def test_driver_writes_the_reference_file_set(tmp_path):
    from tpsc_conductivity_amd.tpsc import CondOpt

    (tmp_path / "input.dat").write_text(f"{U} {tp} {tpp} {t3} {n} {T} {N} {Nw} 1 {mmin}\n")
    out = tmp_path / "out"
    out.mkdir()
    print("[files] tpsc_run 16x128 chijj3=1", flush=True)
    r = subprocess.run([RUN, str(tmp_path / "input.dat"), "--out", str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith(f"U: 4  tp:  -0.175  tpp:  0  t3:  0  n:  0.875  T:  0.1  nk0:  16  Nw0:  128")
    # the same point in-process through the C ABI
    g = CondOpt(N, Nw, n, T, params=(tp, tpp, t3, U), coarse_pass=True)
    g.solve()
    g.calc_chijj_vertex_corr_all_optim(chi3=True, mmin=mmin)
    qx, qy = int(g.scalar("qmax_l")), int(g.scalar("qmax_m"))
    want = names(qx, qy)
    assert sorted(os.listdir(out)) == sorted(want.values())
    # e_k: l, m, eps (precision 15)
    ek = np.array([float(r[2]) for r in table(out / want["ek"], 3)])
    assert np.allclose(ek, g.get("ek"), rtol=1e-14, atol=1e-15)
    # chi0 at (qxmax, qymax) for every n, and at n = 0 for every wedge q
    c = g.get("chi0")
    nk = N // 2 + 1
    nk8 = nk * (nk + 1) // 2
    w = np.array([float(r[1]) for r in table(out / want["chi0w"], 2)])
    assert np.allclose(w, c[qy + qx * (qx + 1) // 2::nk8][: Nw // 2 + 1], rtol=1e-15, atol=0)
    assert np.allclose([float(r[2]) for r in table(out / want["chi0q"], 3)], c[:nk8], rtol=1e-15, atol=0)
    # parameter files
    infos = open(out / want["infos"]).read().splitlines()
    assert infos[0].startswith("Usp:  ") and abs(float(infos[0].split()[1]) - g.scalar("Usp")) < 1e-14
    sp = ["vitesse de Fermi aux points chauds"]
    for kind, r1, r2 in (("(1/(larg.mi-haut.))", "xi_sp/xi_th", "xi_sp*T"),
                         ("(formule asymptotique)", "xi_sp_As/xi_th", "xi_sp_As*T")):
        for ax in ("y", "x"):
            sp += [f"longueur de correlation de spin {kind} (axe {ax})", r1, r2]
    assert [ln.split(":")[0] for ln in infos] == [
        "Usp", "Uch", "chi0max", "double occupation (ansatz)", "qxmax, qymax"] + sp + [
        "U*dblOcc-U*n^2/4", "trace Self*G", "trace Self*G2", "double occupation (Self*G2)", "mu0", "mu"]
    assert abs(float(infos[5].split()[-1]) - g.sp_corr_length()[0]) < 1e-13
    params = [ln.split() for ln in open(out / want["params"])]
    assert [p[0] for p in params] == ["Usp", "Uch", "nbq", "nbw", "dblOcc_ansatz", "qxmax", "qymax",
                                      "chispmax_over_chi0max", "mu0", "mu"]
    assert int(params[2][1]) == nk and int(params[3][1]) == Nw // 2 + 1
    assert abs(float(params[9][1]) - g.scalar("mu")) <= 1e-12 * max(1.0, abs(g.scalar("mu")))
    mu_line = open(out / want["mu"]).read().split()
    assert abs(float(mu_line[1]) - g.scalar("mu")) <= 1e-13 * max(1.0, abs(g.scalar("mu")))
    # local Green function (green.cpp:209-266)
    gl = np.array([complex(float(r[1]), float(r[2])) for r in table(out / want["gloc"], 3)])
    assert np.allclose(gl, g.Gwn_local(), rtol=1e-9, atol=1e-12)
    # Lambda_jj
    assert abs(float(open(out / want["kx0"]).read()) - g.scalar("kx0")) <= 1e-15 * abs(g.scalar("kx0"))
    assert struct.unpack("d", open(out / want["kx0b"], "rb").read())[0] == pytest.approx(g.scalar("kx0"), rel=1e-14)
    j1, j2 = g.get("chijj1"), g.get("chijj2")
    for key, ref in (("jj1", j1), ("jj2", j2), ("jjv", j1 + j2)):
        rows = table(out / want[key], 2)
        assert [int(r[0]) for r in rows] == list(range(Nw // 2 + 1))
        assert "e" in rows[1][1]  # scientific, as the reference's stream
        assert np.allclose([float(r[1]) for r in rows], ref, rtol=1e-14, atol=0)
    # Aslamazov-Larkin on the sparse set
    f = g.get("al_freq").astype(int)
    for key, ref in (("jj3", g.get("chijj3")), ("jja", g.get("chijj_all"))):
        rows = table(out / want[key], 2)
        assert [int(r[0]) for r in rows] == list(f)
        assert np.allclose([float(r[1]) for r in rows], ref, rtol=1e-14, atol=1e-300)
    # Fermi surfaces (green0.cpp:51-105, 626-676; green.cpp:2210-2341) and the MDC (green.cpp:942-1065)
    fs = np.array([[float(x) for x in r] for r in table(out / want["fs"], 2)])
    assert np.abs(fs - g.find_FS(10001)).max() < 1e-9  # setprecision(10)
    nk8 = (N // 2 + 1) * (N // 2 + 2) // 2
    m0, a0 = g.find_FS_PBC()
    rows = table(out / want["fsdA"], 3)
    assert [int(r[1]) for r in rows] == list(m0) and [int(r[0]) for r in rows] == [l + m for l, m in enumerate(m0)]
    assert [int(r[1]) for r in table(out / want["fsd"], 2)] == list(m0)
    (mp, ap), (mq, aq) = g.find_FS_inter()
    assert [int(r[1]) for r in table(out / want["fsiA"], 3)] == list(mp)
    assert [int(r[1]) for r in table(out / want["fsiAP"], 3)] == list(mq)
    assert len(table(out / want["fsi"], 2)) == len(table(out / want["fsiP"], 2)) == N // 2 + 1
    sw, ak = g.MDC(list(range(Nw // 2)), 10, 0.0, [0.01, 0.001, 0.0001, 0.00001])
    rows = table(out / want["mdcS"], 2 + 8)
    assert len(rows) == nk8
    assert np.allclose([float(r[2]) for r in rows], sw[:, 0].real, rtol=1e-11, atol=1e-13)
    rows = table(out / want["mdcA"], 2 + 4)
    assert np.allclose([[float(x) for x in r[2:]] for r in rows], ak, rtol=1e-11, atol=1e-13)
    # angle-binned Lambda_jj with one bin (cond_opt.cpp:3889-3946)
    rows = table(out / want["jj1kt"], 4)
    assert np.allclose([float(r[1]) for r in rows], j1, rtol=1e-14, atol=0)
    assert [int(r[1]) for r in table(out / want["jj2k"], 3)] == [1] * (Nw // 2 + 1)
    g.close()


def test_write_files_follows_the_completed_stages(tmp_path):
    from tpsc_conductivity_amd.tpsc import CondOpt

    g = CondOpt(8, 32, 0.9, 0.25, params=(-0.1, 0.0, 0.0, 2.5), coarse_pass=False)
    assert g.write_files(str(tmp_path)) == 4           # e_k, the Fermi surfaces of the constructor
    g.find_Usp_Uch()
    assert g.write_files(str(tmp_path)) == 8           # + chi0 files, TPSC_infos, TPSCparams
    assert not any(p.startswith("mu0_mu_D") for p in os.listdir(tmp_path))
    g.close()
