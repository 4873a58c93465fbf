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
