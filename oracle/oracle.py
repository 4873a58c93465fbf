"""ctypes wrapper of the FP64 CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
and only as the checker.  The oracle restates dbergeron1/TPSC_conductivity's hot path
(see oracle/tpsc_oracle.h for the file:line map).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libtpsc_oracle.so")


def build(force=False):
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def use_native():
    """Build the oracle with -march=native for this host (oracle/_build_native) and load that copy
    instead of the portable one; must run before the first lib() call.  Returns its path."""
    global _LIB_PATH
    if _lib is not None:
        raise RuntimeError("oracle library already loaded")
    subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True)
    _LIB_PATH = os.path.join(_HERE, "_build_native", "libtpsc_oracle.so")
    return _LIB_PATH


class OrcParams(C.Structure):
    _fields_ = [
        ("nk0", C.c_int), ("Nw0", C.c_int), ("n", C.c_double), ("T", C.c_double),
        ("eta", C.c_double), ("tp", C.c_double), ("tpp", C.c_double), ("t3", C.c_double),
        ("U", C.c_double), ("coarse_pass", C.c_int), ("t", C.c_double),
    ]


_dp = C.POINTER(C.c_double)


class Orc(C.Structure):
    _fields_ = [
        ("p", OrcParams),
        ("nbk", C.c_int), ("nbw", C.c_int), ("nwmax", C.c_int),
        ("ek", _dp), ("er", _dp), ("mu0", C.c_double),
        ("chiqw", _dp), ("chiqt", _dp), ("chirt", _dp),
        ("dtau_chirt0", _dp), ("dtau_Grt0p", _dp), ("dtau_chiqt0", _dp),
        ("chi0_inf", _dp), ("chi0_inf2", _dp),
        ("chi0max", C.c_double), ("qmax", C.c_int * 2),
        ("nsr", C.c_double), ("Usp", C.c_double), ("Uch", C.c_double), ("dblOcc", C.c_double),
        ("difRelsp", C.c_double), ("difRelch", C.c_double),
        ("Usp_coarse", C.c_double), ("Uch_coarse", C.c_double),
        ("nevals_sp", C.c_int), ("nevals_ch", C.c_int),
        ("V", _dp),
        ("Sigma_tau", _dp), ("dtau_Sigma", _dp), ("Self", _dp),
        ("Sigma_inf", C.c_double), ("Sigma_inf2", _dp), ("Sigma_inf3", _dp), ("Sigma_inf4", _dp),
        ("trSG", C.c_double), ("trSG_corr", C.c_double), ("trSG2", C.c_double), ("dblOcc2", C.c_double),
        ("mu", C.c_double), ("mu_ok", C.c_int), ("nevals_mu", C.c_int),
        ("chijj1", _dp), ("chijj2", _dp), ("kx0", C.c_double),
        ("chijj1_inf2", C.c_double), ("chijj2_inf2", C.c_double), ("chijj2_inf4", C.c_double),
        ("N_theta", C.c_int),
        ("chijj1_k", _dp), ("chijj2_k", _dp),
        ("al_N", C.c_int), ("al_N_theta", C.c_int), ("al_freq", C.POINTER(C.c_long)),
        ("chijj3", _dp), ("chijj_all", _dp), ("chijj3_k", _dp), ("al_hw", _dp), ("al_hg2", _dp),
        ("mu0_fixed", C.c_int),
    ]


_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double)
_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER(Orc)
        L.orc_create.restype = P
        L.orc_create.argtypes = [C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [P]
        L.orc_solve_timed.argtypes = [P, _dp]
        L.orc_solve_timed.restype = C.c_int
        L.orc_omp_max_threads.restype = C.c_int
        for name in ("orc_find_mu0", "orc_find_Usp_Uch", "orc_vertex", "orc_self_energy",
                     "orc_trace_self_G", "orc_find_mu", "orc_trace_self_G2", "orc_solve"):
            getattr(L, name).argtypes = [P]
            getattr(L, name).restype = C.c_int
        L.orc_set_grid.argtypes = [P, C.c_int, C.c_int]
        L.orc_chi0.argtypes = [P, C.c_int, C.c_int]
        L.orc_chi0.restype = C.c_int
        L.orc_chijj.argtypes = [P, C.c_int]
        L.orc_chijj.restype = C.c_int
        L.orc_fix_mu0.argtypes = [P, C.c_double]
        L.orc_solve_from_roots.argtypes = [P, C.c_double, C.c_double]
        L.orc_solve_from_roots.restype = C.c_int
        L.orc_hot_spot_vF.argtypes = [P]
        L.orc_hot_spot_vF.restype = C.c_double
        L.orc_sp_corr_from.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _dp]
        L.orc_sp_corr_length.argtypes = [P, _dp]
        L.orc_sp_corr_length.restype = C.c_int
        L.orc_sigma_loss.argtypes = [P, _dp]
        L.orc_sigma_loss.restype = C.c_double
        L.orc_loss.argtypes = [P, _dp, C.c_int]
        L.orc_loss.restype = C.c_double
        L.orc_gwn_local.argtypes = [P, _dp]
        L.orc_gwn_local.restype = C.c_int
        L.orc_chijj_al.argtypes = [P, C.c_int, C.c_int]
        L.orc_chijj_al.restype = C.c_int
        L.orc_al_freqs.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_long)]
        L.orc_al_freqs.restype = C.c_int
        L.orc_sum_rule.argtypes = [P, C.c_int, C.c_double]
        L.orc_sum_rule.restype = C.c_double
        L.orc_find_Usp_Uch_grid.argtypes = [P, C.c_int, C.c_int, _dp]
        L.orc_find_Usp_Uch_grid.restype = C.c_int
        L.orc_traceDiffG.argtypes = [P, C.c_double]
        L.orc_traceDiffG.restype = C.c_double
        L.orc_ddensity_mu_PBC.argtypes = [P, C.c_double]
        L.orc_ddensity_mu_PBC.restype = C.c_double
        L.orc_chiqw_discret.argtypes = [P, C.c_int, C.c_int, C.c_int]
        L.orc_chiqw_discret.restype = C.c_double
        L.orc_trace_chi0.argtypes = [P, _dp]
        L.orc_trace_chi0.restype = C.c_double
        L.orc_self_kw_brute_out.argtypes = [P, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_find_zero.argtypes = [_FN, C.c_void_p, _dp, _dp, _dp, C.POINTER(C.c_int)]
        L.orc_find_zero.restype = C.c_int
        L.orc_spline_coeffs_rel.argtypes = [_dp, _dp, C.c_int, _dp]
        L.orc_roots_2pol.argtypes = [_dp, _dp]
        L.orc_pade_coef.argtypes = [_dp, _dp, C.c_int, _dp]
        L.orc_pade_eval_out.argtypes = [_dp, C.c_int, _dp, _dp, _dp]
        L.orc_r2r_2d.argtypes = [_dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_redft00_1d.argtypes = [_dp, C.c_int]
        L.orc_dft.argtypes = [_dp, C.c_int, C.c_int]
        L.orc_dispk.argtypes = [P, C.c_double, C.c_double]
        L.orc_dispk.restype = C.c_double
        L.orc_grad_ek_x.argtypes = [P, C.c_double, C.c_double]
        L.orc_grad_ek_x.restype = C.c_double
        L.orc_al_tail_out.argtypes = [C.c_double] * 5 + [C.c_int, _dp]
        _lib = L
    return _lib


def omp_max_threads():
    return lib().orc_omp_max_threads()


def _arr(ptr, n, dtype=np.float64):
    if not ptr:
        return None
    a = np.ctypeslib.as_array(ptr, shape=(n,))
    return a.copy() if dtype == np.float64 else a.copy().view(np.complex128)


class Oracle:
    """One parameter point of the TPSC chain, mirroring the reference's object methods."""

    def __init__(self, N, Nw, n, T, U, tp=0.0, tpp=0.0, t3=0.0, eta=0.1, coarse_pass=True, t=1.0, mu0=None):
        self.params = OrcParams(N, Nw, n, T, eta, tp, tpp, t3, U, int(coarse_pass), t)
        self._o = lib().orc_create(C.byref(self.params))
        if mu0 is not None:
            lib().orc_fix_mu0(self._o, mu0)

    def __del__(self):
        try:
            if self._o:
                lib().orc_destroy(self._o)
                self._o = None
        except Exception:
            pass

    @property
    def s(self):
        return self._o.contents

    # ---- sizes
    @property
    def nk(self):
        return self.s.nbk

    @property
    def nk8(self):
        return self.s.nbk * (self.s.nbk + 1) // 2

    @property
    def nwb(self):
        return self.s.nwmax + 1

    @property
    def nq(self):
        return self.s.nbk * (self.s.nbk - 2)

    # ---- stages
    def find_mu0(self):
        return lib().orc_find_mu0(self._o)

    def calc_chiqw_FFT_spline(self, nq0, nw0):
        return lib().orc_chi0(self._o, nq0, nw0)

    def sum_rule(self, channel, Ux):
        return lib().orc_sum_rule(self._o, channel, Ux)

    def find_Usp_Uch(self):
        return lib().orc_find_Usp_Uch(self._o)

    def calc_chirt(self):
        return lib().orc_vertex(self._o)

    def calc_Self(self):
        return lib().orc_self_energy(self._o)

    def calc_Self_FFT_spline(self):
        """green.cpp:1673: sum rules, vertex, then the self-energy."""
        self.find_Usp_Uch()
        self.calc_chirt()
        return self.calc_Self()

    def traceSelfG(self):
        return lib().orc_trace_self_G(self._o)

    def traceDiffG(self, mu):
        return lib().orc_traceDiffG(self._o, mu)

    def find_mu(self):
        return lib().orc_find_mu(self._o)

    def calc_chijj(self, N_theta=1):
        return lib().orc_chijj(self._o, N_theta)

    def calc_chijj_al(self, mmin, N_theta=1):
        """Aslamazov-Larkin Lambda_jj3 on the sparse boson set (cond_opt.cpp:4292-6467)."""
        return lib().orc_chijj_al(self._o, mmin, N_theta)

    def sigma_loss(self, w):
        """sum_k w_k Re Sigma(k, i w_0) after chi0 -> Usp, Uch -> V -> Sigma (NaN if a root fails)."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.size == self.nk8
        return lib().orc_sigma_loss(self._o, w.ctypes.data_as(_dp))

    def loss(self, w, terms):
        """The loss terms of tpsc_loss_forward (bit 1 Re Sigma(k, i pi T), 2 Im, 4 chi_sp(pi,pi,0),
        8 sum_{n <= 8} Lambda_jj(iq_n)) after the chain they need; NaN if a root fails."""
        if w is None:
            w = np.zeros(self.nk8)
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.size == self.nk8
        return lib().orc_loss(self._o, w.ctypes.data_as(_dp), int(terms))

    def sp_corr_length(self):
        """(vF, xi_y, xi_x, xi_as_y, xi_as_x) of chi::sp_corr_length (chi.cpp:451-640)."""
        out = np.zeros(5)
        assert lib().orc_sp_corr_length(self._o, out.ctypes.data_as(_dp)) == 0
        return out

    def Gwn_local(self):
        """G_loc(iw_n), n < Nw/2 (green.cpp:209-266)."""
        out = np.zeros(self.s.nbw // 2, dtype=np.complex128)
        assert lib().orc_gwn_local(self._o, out.ctypes.data_as(_dp)) == 0
        return out

    def al_freqs(self):
        s = self.s
        return np.array([s.al_freq[i] for i in range(s.al_N)], dtype=np.int64)

    def solve(self):
        return lib().orc_solve(self._o)

    def solve_from_roots(self, Usp, Uch=float("nan")):
        """After solve(): the chain again from the given Usp and Uch (Uch NaN: its root searched
        at that Usp), then vertex, self-energy, mu and Lambda_jj, on the same chi0."""
        return lib().orc_solve_from_roots(self._o, Usp, Uch)

    def solve_timed(self):
        """A full solve without the coarse pass; (ok, seconds of the stages A mu0 + chi0, B roots,
        C vertex, D self-energy, E traceSelfG + mu, F Lambda_jj)."""
        t = np.zeros(6)
        ok = lib().orc_solve_timed(self._o, t.ctypes.data_as(_dp))
        return bool(ok), t

    def chiqw_discret(self, qx, qy, n):
        return lib().orc_chiqw_discret(self._o, qx, qy, n)

    def trace_chi0(self):
        corr = C.c_double(0.0)
        tr = lib().orc_trace_chi0(self._o, C.byref(corr))
        return tr, corr.value

    def dispersion(self, kx, ky):
        return lib().orc_dispk(self._o, kx, ky)

    def velocity_x(self, kx, ky):
        return lib().orc_grad_ek_x(self._o, kx, ky)

    def self_kw_brute(self, kx, ky, n):
        r = (C.c_double * 2)()
        lib().orc_self_kw_brute_out(self._o, kx, ky, n, r)
        return complex(r[0], r[1])

    # ---- arrays (copies)
    def get(self, name):
        s = self.s
        nk8, nwb, nk = self.nk8, self.nwb, self.nk
        sizes = {
            "ek": nk * nk, "er": nk8, "chiqw": nwb * nk8, "chiqt": nwb * nk8, "chirt": nwb * nk8,
            "dtau_chirt0": nk8, "dtau_Grt0p": nk8, "dtau_chiqt0": nk8, "chi0_inf": nk8,
            "chi0_inf2": nk8, "V": nwb * nk8, "Sigma_tau": (s.nbw + 1) * nk8, "dtau_Sigma": 2 * nk8,
            "Sigma_inf2": nk8, "Sigma_inf3": nk8, "Sigma_inf4": nk8, "chijj1": nwb, "chijj2": nwb,
            "chijj1_k": nwb * s.N_theta, "chijj2_k": nwb * s.N_theta,
            "chijj3": s.al_N, "chijj_all": s.al_N, "chijj3_k": s.al_N * s.al_N_theta,
            "al_hw": self.nq * (s.nbw // 2), "al_hg2": self.nq,
        }
        if name == "Self":
            return _arr(s.Self, 2 * nwb * nk8, np.complex128)
        return _arr(getattr(s, name), sizes[name])

    def scalars(self):
        s = self.s
        keys = ["mu0", "chi0max", "nsr", "Usp", "Uch", "dblOcc", "difRelsp", "difRelch", "Usp_coarse",
                "Uch_coarse", "Sigma_inf", "trSG", "trSG_corr", "trSG2", "dblOcc2", "mu", "mu_ok",
                "kx0", "chijj1_inf2", "chijj2_inf2", "chijj2_inf4", "nevals_sp", "nevals_ch", "nevals_mu"]
        d = {k: getattr(s, k) for k in keys}
        d["qmax"] = (s.qmax[0], s.qmax[1])
        return d


# ---- numerics helpers for tests -------------------------------------------------------
def r2r_2d(a, kind0, kind1):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    lib().orc_r2r_2d(a.ctypes.data_as(_dp), a.shape[0], a.shape[1], kind0, kind1)
    return a


def redft00(a):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    lib().orc_redft00_1d(a.ctypes.data_as(_dp), a.size)
    return a


def dft(a, sign):
    a = np.ascontiguousarray(a, dtype=np.complex128).copy()
    lib().orc_dft(a.ctypes.data_as(_dp), a.size, sign)
    return a


def sp_corr_from(c0, nbq, qx, qy, chi0max, Usp):
    c0 = np.ascontiguousarray(c0, dtype=np.float64)
    out = np.zeros(4)
    lib().orc_sp_corr_from(c0.ctypes.data_as(_dp), nbq, qx, qy, chi0max, Usp, out.ctypes.data_as(_dp))
    return out


def al_freqs(nbw, mmin):
    n = lib().orc_al_freqs(nbw, mmin, None)
    out = (C.c_long * n)()
    lib().orc_al_freqs(nbw, mmin, out)
    return np.array(list(out), dtype=np.int64)


def al_tail(Usp, Uch, chi0inf, T, tau, kind):
    """Imaginary-time two-pole tail of the AL vertex (oracle orc_al_tail)."""
    out = (C.c_double * 2)()
    lib().orc_al_tail_out(Usp, Uch, chi0inf, T, tau, kind, out)
    return complex(out[0], out[1])


def spline_coeffs_rel(x, v, d0, dN):
    x = np.ascontiguousarray(x, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    c = np.zeros(4 * (len(x) - 1))
    c[0], c[1] = d0, dN
    lib().orc_spline_coeffs_rel(x.ctypes.data_as(_dp), v.ctypes.data_as(_dp), len(x), c.ctypes.data_as(_dp))
    return c.reshape(-1, 4)


def find_zero(f, init, lims=None):
    cb = _FN(lambda ctx, x: f(x))
    init = (C.c_double * 3)(*init)
    root = C.c_double(0.0)
    ne = C.c_int(0)
    lp = (C.c_double * 2)(*lims) if lims is not None else None
    ok = lib().orc_find_zero(cb, None, C.cast(init, _dp), C.byref(root),
                             C.cast(lp, _dp) if lp is not None else None, C.byref(ne))
    return bool(ok), root.value, ne.value


def roots_2pol(coef):
    c = np.ascontiguousarray(coef, dtype=np.complex128)
    r = np.zeros(2, dtype=np.complex128)
    lib().orc_roots_2pol(c.ctypes.data_as(_dp), r.ctypes.data_as(_dp))
    return r


def pade(func, x, z):
    func = np.ascontiguousarray(func, dtype=np.complex128)
    x = np.ascontiguousarray(x, dtype=np.complex128)
    coef = np.zeros_like(func)
    L = lib()
    L.orc_pade_coef(func.ctypes.data_as(_dp), x.ctypes.data_as(_dp), len(func), coef.ctypes.data_as(_dp))
    out = []
    for zz in np.atleast_1d(z):
        zi = (C.c_double * 2)(zz.real, zz.imag)
        r = (C.c_double * 2)()
        L.orc_pade_eval_out(zi, len(func), x.ctypes.data_as(_dp), coef.ctypes.data_as(_dp), r)
        out.append(complex(r[0], r[1]))
    return np.array(out)
