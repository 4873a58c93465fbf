"""Host-side mirror of the reference's object interface for the TPSC hot path.

`CondOpt` keeps the method names, argument meaning and call order of the reference class
chain cond_opt > green > chi > chi0 > green0 (dbergeron1/TPSC_conductivity); each method is
one C-ABI call into libtpsc_hip.so, where all the arrays live on the GPU.  Errors that the
reference reports on stdout (a root not found, a stage called too early) raise TPSCError.
"""
import ctypes as C

from .capi import PRECISIONS, Context, TPSCError, lib, pade_conductivity

__all__ = ["CondOpt", "TPSCError"]


class CondOpt:
    """cond_opt(nk0, Nw0, n, T, eta, params={t', t'', t3, U}) -- cond_opt.h:25.

    precision: "fp64" (the reference's arithmetic) or "mixed" (FP32 slab transforms and
    tau/iqn arrays, FP64 boundary data, moments, roots and Matsubara sums).
    """

    def __init__(self, nk0, Nw0, n, T, eta=0.1, params=(0.0, 0.0, 0.0, 1.0), coarse_pass=True, precision="fp64",
                 t=1.0):
        tp, tpp, t3, U = params
        self.precision = precision
        self.ctx = Context(nk0, Nw0, n, T, U, tp=tp, tpp=tpp, t3=t3, eta=eta, coarse_pass=coarse_pass,
                           precision=PRECISIONS[precision], t=t)
        self.nk0, self.Nw0, self.density, self.tem, self.eta = nk0, Nw0, n, T, eta
        self.tp, self.tpp, self.t3, self.U = tp, tpp, t3, U

    def close(self):
        self.ctx.close()

    # green0.cpp:240-262
    def find_mu0(self):
        v = C.c_double(0.0)
        self.ctx.check(lib().tpsc_find_mu0(self.ctx.handle, C.byref(v)), "find_mu0")
        return v.value

    # chi0.cpp:279
    def calc_chiqw_FFT_spline(self, nq0, nw0):
        self.ctx.check(lib().tpsc_chi0(self.ctx.handle, nq0, nw0), "calc_chiqw_FFT_spline")

    # chi.cpp:833-843
    def sum_rule_sp(self, Us):
        v = C.c_double(0.0)
        self.ctx.check(lib().tpsc_sum_rule(self.ctx.handle, 0, Us, C.byref(v)), "sum_rule_sp")
        return v.value

    def sum_rule_ch(self, Uc):
        v = C.c_double(0.0)
        self.ctx.check(lib().tpsc_sum_rule(self.ctx.handle, 1, Uc, C.byref(v)), "sum_rule_ch")
        return v.value

    # chi.cpp:210 / :846
    def find_Usp_Uch(self, nq0=None, nw0=None):
        a, b = C.c_double(0.0), C.c_double(0.0)
        if nq0 is None:
            rc = lib().tpsc_find_Usp_Uch(self.ctx.handle, C.byref(a), C.byref(b))
        else:
            rc = lib().tpsc_find_Usp_Uch_grid(self.ctx.handle, nq0, nw0, C.byref(a), C.byref(b))
        self.ctx.check(rc, "find_Usp_Uch")
        return a.value, b.value

    # chi.cpp:37
    def calc_chirt(self):
        self.ctx.check(lib().tpsc_vertex_rt(self.ctx.handle), "calc_chirt")

    # green.cpp:1673
    def calc_Self_FFT_spline(self):
        self.ctx.check(lib().tpsc_calc_self_energy_full(self.ctx.handle), "calc_Self_FFT_spline")

    def calc_Self(self):
        """Sigma from an existing vertex (green.cpp:1737-2031 without the leading stages)."""
        self.ctx.check(lib().tpsc_self_energy(self.ctx.handle), "calc_Self")

    # green.cpp:3001
    def traceSelfG(self):
        a, b = C.c_double(0.0), C.c_double(0.0)
        self.ctx.check(lib().tpsc_trace_self_G(self.ctx.handle, C.byref(a), C.byref(b)), "traceSelfG")
        return a.value, b.value

    # green.cpp:3108
    def traceDiffG(self, mu):
        v = C.c_double(0.0)
        self.ctx.check(lib().tpsc_trace_diff_G(self.ctx.handle, mu, C.byref(v)), "traceDiffG")
        return v.value

    # green.cpp:3185
    def find_mu(self):
        v = C.c_double(0.0)
        self.ctx.check(lib().tpsc_find_mu(self.ctx.handle, C.byref(v)), "find_mu")
        return v.value

    # cond_opt.cpp:2537
    def calc_chijj_vertex_corr_all_optim(self, chi3=False, mmin=3, N_theta_k=1, k_chijj=None, nk_chijj=0):
        self.ctx.check(lib().tpsc_chijj(self.ctx.handle, int(bool(chi3)), mmin, N_theta_k),
                       "calc_chijj_vertex_corr_all_optim")

    def solve(self):
        self.ctx.check(lib().tpsc_solve(self.ctx.handle), "solve")

    # results
    def get(self, name):
        return self.ctx.get(name)

    def scalar(self, name):
        return self.ctx.scalar(name)

    def conductivity_pade(self, omega, eta=None, npade=None):
        """sigma(w) = Im Lambda_jj(w + i eta)/w by the bosonic Pade continuation of
        Lambda_jj1 + Lambda_jj2 (chi.cpp:1055-1093 applied to Lambda_jj); eta defaults to the
        constructor's eta, npade to all Nw/2 + 1 Matsubara points."""
        lam = self.get("chijj1") + self.get("chijj2")
        return pade_conductivity(lam, self.tem, len(lam) if npade is None else npade, omega,
                                 self.eta if eta is None else eta)

    def Gwn_local(self):
        """green::Gwn_local (green.cpp:209-266): G_loc(iw_n) for n < Nw/2 (complex array)."""
        import numpy as np

        nw = int(self.scalar("nwb")) - 1
        out = np.zeros(nw, dtype=np.complex128)
        self.ctx.check(lib().tpsc_gwn_local(self.ctx.handle, out.ctypes.data_as(C.POINTER(C.c_double))),
                       "Gwn_local")
        return out

    def sp_corr_length(self):
        """chi::sp_corr_length (chi.cpp:451-640): (vF, xi_y, xi_x, xi_as_y, xi_as_x)."""
        import numpy as np

        out = np.zeros(5)
        self.ctx.check(lib().tpsc_sp_corr_length(self.ctx.handle, out.ctypes.data_as(C.POINTER(C.c_double))),
                       "sp_corr_length")
        return out

    def fix_mu0(self, mu0):
        """Make mu0 an input (every later stage uses it instead of solving the density); NaN releases."""
        self.ctx.check(lib().tpsc_fix_mu0(self.ctx.handle, float(mu0)), "tpsc_fix_mu0")

    def sigma_loss_forward(self, w):
        """L = sum_k w_k Re Sigma(k, i w_0) after calc_Self_FFT_spline (FP64 path)."""
        import numpy as np

        w = np.ascontiguousarray(w, dtype=np.float64)
        L = C.c_double(0.0)
        self.ctx.check(lib().tpsc_sigma_loss_forward(self.ctx.handle, w.ctypes.data_as(C.POINTER(C.c_double)),
                                                     C.byref(L)), "tpsc_sigma_loss_forward")
        return L.value

    def sigma_loss_backward(self, dloss=1.0):
        """dloss * dL/d(t, t', mu0, U) for the last sigma_loss_forward."""
        import numpy as np

        g = np.zeros(4)
        self.ctx.check(lib().tpsc_sigma_loss_backward(self.ctx.handle, float(dloss),
                                                      g.ctypes.data_as(C.POINTER(C.c_double))),
                       "tpsc_sigma_loss_backward")
        return g

    def loss_forward(self, w, terms):
        """L = the loss terms `terms` (capi.LOSS_*: Re / Im Sigma(k, i pi T) weighted by w,
        chi_sp(pi, pi, 0), sum_{n <= 8} Lambda_jj(iq_n)) after the chain they need (FP64 path)."""
        import numpy as np

        wp = None
        if w is not None:
            w = np.ascontiguousarray(w, dtype=np.float64)
            wp = w.ctypes.data_as(C.POINTER(C.c_double))
        L = C.c_double(0.0)
        self.ctx.check(lib().tpsc_loss_forward(self.ctx.handle, wp, int(terms), C.byref(L)), "tpsc_loss_forward")
        return L.value

    def loss_backward(self, dloss=1.0):
        """dloss * dL/d(t, t', mu0, U) for the last loss_forward."""
        import numpy as np

        g = np.zeros(4)
        self.ctx.check(lib().tpsc_loss_backward(self.ctx.handle, float(dloss), g.ctypes.data_as(C.POINTER(C.c_double))),
                       "tpsc_loss_backward")
        return g

    # ---- post-processing (host FP64 over the device results, tpsc_main.cpp:77-101)
    def find_mu_inf_syst(self):
        """green0::find_mu_inf_syst (green0.cpp:266-287): mu0 of the infinite lattice."""
        v = C.c_double(0.0)
        self.ctx.check(lib().tpsc_mu0_infinite(self.ctx.handle, C.byref(v)), "find_mu_inf_syst")
        return v.value

    def find_FS(self, nk=10001):
        """green0::find_FS (green0.cpp:626-676): (kx, ky) on the non-interacting Fermi surface along
        kx - ky = j pi/(nk-1), as an (count, 2) array."""
        import numpy as np

        out = np.zeros(2 * nk)
        cnt = C.c_int(0)
        self.ctx.check(lib().tpsc_fermi_surface(self.ctx.handle, nk, out.ctypes.data_as(C.POINTER(C.c_double)),
                                                C.byref(cnt)), "find_FS")
        return out[: 2 * cnt.value].reshape(-1, 2)

    def _fs_lattice(self, kind, what):
        import numpy as np

        nk = self.nk0 // 2 + 1
        m = np.zeros(nk, dtype=np.int32)
        a = np.zeros(nk)
        self.ctx.check(lib().tpsc_fs_lattice(self.ctx.handle, kind, m.ctypes.data_as(C.POINTER(C.c_int)),
                                             a.ctypes.data_as(C.POINTER(C.c_double))), what)
        return m, a

    def find_FS_PBC(self):
        """green0::find_FS_PBC (green0.cpp:51-105): per diagonal l, the m of largest A0(k, 0) and that A0."""
        return self._fs_lattice(0, "find_FS_PBC")

    def find_FS_inter(self):
        """green::find_FS_inter (green.cpp:2210-2341): ((m, A) from the 3-point polynomial, (m, A) from the
        4-point Pade) per diagonal l."""
        return self._fs_lattice(1, "find_FS_inter"), self._fs_lattice(2, "find_FS_inter")

    def MDC(self, n, NP, w, eta_p, use_Pade=True):
        """green::MDC (green.cpp:814-1068): (Sigma(k, w + i eta) [nk8, Neta] complex, A [nk8, Neta])."""
        import numpy as np

        nn = np.ascontiguousarray(n[:max(NP, 2)], dtype=np.int32)
        eta = np.ascontiguousarray(eta_p, dtype=np.float64)
        nk8 = int(self.scalar("nk8"))
        sw = np.zeros((nk8, len(eta)), dtype=np.complex128)
        ak = np.zeros((nk8, len(eta)))
        dp = C.POINTER(C.c_double)
        self.ctx.check(lib().tpsc_mdc(self.ctx.handle, nn.ctypes.data_as(C.POINTER(C.c_int)), NP, w,
                                      eta.ctypes.data_as(dp), len(eta), int(bool(use_Pade)), sw.ctypes.data_as(dp),
                                      ak.ctypes.data_as(dp)), "MDC")
        return sw, ak

    def EDC(self, kx, ky, w, m_freq, NP_max, eta_p):
        """green::EDC(kx, ky, Nk, w, nw0, m_freq, NP_max, eta, N_eta) (green.cpp:589-700): the spectral
        weight [point, w, eta]."""
        import numpy as np

        kx = np.ascontiguousarray(kx, dtype=np.int32)
        ky = np.ascontiguousarray(ky, dtype=np.int32)
        w = np.ascontiguousarray(w, dtype=np.float64)
        eta = np.ascontiguousarray(eta_p, dtype=np.float64)
        out = np.zeros((len(kx), len(w), len(eta)))
        ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
        self.ctx.check(lib().tpsc_edc(self.ctx.handle, kx.ctypes.data_as(ip), ky.ctypes.data_as(ip), len(kx),
                                      w.ctypes.data_as(dp), len(w), m_freq, NP_max, eta.ctypes.data_as(dp), len(eta),
                                      out.ctypes.data_as(dp)), "EDC")
        return out

    def calc_Self_Re_w(self, k, n, w, eta_p):
        """green::calc_Self_Re_w (green.cpp:408-551): (Sigma, A, Sigma_Jfrac, A_Jfrac), each [w, eta]."""
        import numpy as np

        kk = np.ascontiguousarray(k, dtype=np.int32)
        nn = np.ascontiguousarray(n, dtype=np.int32)
        w = np.ascontiguousarray(w, dtype=np.float64)
        eta = np.ascontiguousarray(eta_p, dtype=np.float64)
        s1 = np.zeros((len(w), len(eta)), dtype=np.complex128)
        s2 = np.zeros_like(s1)
        a1 = np.zeros((len(w), len(eta)))
        a2 = np.zeros_like(a1)
        ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
        self.ctx.check(lib().tpsc_self_real_freq(self.ctx.handle, kk.ctypes.data_as(ip), nn.ctypes.data_as(ip),
                                                 len(nn), w.ctypes.data_as(dp), len(w), eta.ctypes.data_as(dp),
                                                 len(eta), s1.ctypes.data_as(dp), a1.ctypes.data_as(dp),
                                                 s2.ctypes.data_as(dp), a2.ctypes.data_as(dp)), "calc_Self_Re_w")
        return s1, a1, s2, a2

    def save_Self_bin(self, directory=None):
        """green::save_Self_bin (green.cpp:1496-1514), FP64 path."""
        self.ctx.check(lib().tpsc_save_self_bin(self.ctx.handle, directory.encode() if directory else None),
                       "save_Self_bin")

    def save_chi(self, directory=None):
        """chi0::save_chi (chi0.cpp:61-96)."""
        self.ctx.check(lib().tpsc_save_chi0_bin(self.ctx.handle, directory.encode() if directory else None),
                       "save_chi")

    def write_files(self, directory=None):
        """Write the reference's text outputs (TPSCparams_*, TPSC_infos_*, chijj*_qn_*, kx0_*, ...)
        for the stages run so far into directory (default: the current directory)."""
        return self.ctx.write_files(directory)

    def keep_chi_tau(self, on=True):
        self.ctx.check(lib().tpsc_keep_chi_tau(self.ctx.handle, int(on)), "keep_chi_tau")

    def profile(self, on=True):
        self.ctx.check(lib().tpsc_prof_enable(self.ctx.handle, int(on)), "prof_enable")

    def profile_reset(self):
        self.ctx.check(lib().tpsc_prof_reset(self.ctx.handle), "prof_reset")

    def profile_table(self):
        L = lib()
        out = {}
        for i in range(L.tpsc_prof_count(self.ctx.handle)):
            name, ms, calls = C.c_char_p(), C.c_double(), C.c_int()
            self.ctx.check(L.tpsc_prof_get(self.ctx.handle, i, C.byref(name), C.byref(ms), C.byref(calls)), "prof_get")
            out[name.value.decode()] = (ms.value, calls.value)
        return out
