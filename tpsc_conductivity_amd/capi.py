"""ctypes binding of libtpsc_hip.so (the C ABI declared in include/tpsc_capi.h).

The product path: every computation runs in the HIP library.  If the shared library is
missing this module raises at import time instead of falling back to anything else.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libtpsc_hip.so")

TPSC_OK = 0
ERRORS = {-1: "TPSC_E_INVALID", -2: "TPSC_E_DEVICE", -3: "TPSC_E_NOMEM", -4: "TPSC_E_ORDER", -5: "TPSC_E_NOROOT"}

FIELDS = {
    "ek": 1, "chi0": 2, "chirt": 3, "chi0_inf": 4, "chi0_inf2": 5, "dtau_chirt0": 6, "V": 7,
    "Sigma_tau": 8, "Self": 9, "Sigma_inf2": 10, "Sigma_inf3": 11, "Sigma_inf4": 12,
    "chijj1": 13, "chijj2": 14, "dtau_Sigma": 15, "chi0_tau": 16, "Sigma_kw": 17,
    "chijj3": 18, "chijj_all": 19, "al_freq": 20, "chijj1_k": 21, "chijj2_k": 22, "chijj3_k": 23,
    "grad_terms": 24, "grad_dmu": 25, "grad_dlambda": 26, "Sigma_bnd": 27,
}
# loss terms of tpsc_loss_forward (bitwise or)
LOSS_RE_SIGMA, LOSS_IM_SIGMA, LOSS_CHISP_Q, LOSS_LAMBDA = 1, 2, 4, 8
PREC_FP64 = 0
PREC_MIXED = 1
PRECISIONS = {"fp64": PREC_FP64, "mixed": PREC_MIXED}
SCALARS = {
    "mu0": 1, "chi0max": 2, "qmax_l": 3, "qmax_m": 4, "Usp": 5, "Uch": 6, "dblOcc": 7,
    "difRelsp": 8, "difRelch": 9, "Sigma_inf": 10, "trSG": 11, "trSG_corr": 12, "mu": 13,
    "trSG2": 14, "dblOcc2": 15, "kx0": 16, "chijj1_inf2": 17, "chijj2_inf2": 18,
    "chijj2_inf4": 19, "nevals_sp": 20, "nevals_ch": 21, "nevals_mu": 22, "nk": 23, "nk8": 24,
    "nwb": 25, "nq": 26, "al_N": 27, "al_mmin": 28, "stages": 29, "n_theta": 30, "chi0_fp64": 31,
    "roots_on_device": 32,
}


class TPSCError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("nk0", C.c_int), ("Nw0", C.c_int), ("n", C.c_double), ("T", C.c_double),
        ("eta", C.c_double), ("tp", C.c_double), ("tpp", C.c_double), ("t3", C.c_double),
        ("U", C.c_double), ("coarse_pass", C.c_int), ("precision", C.c_int), ("t", C.c_double),
    ]


class BatchResult(C.Structure):
    _fields_ = [("status", C.c_int), ("queued", C.c_int), ("mu0", C.c_double), ("Usp", C.c_double), ("Uch", C.c_double),
                ("mu", C.c_double), ("dblOcc", C.c_double), ("kx0", C.c_double), ("seconds", C.c_double)]


_lib = None

EXPORTS = [
    "tpsc_create", "tpsc_destroy", "tpsc_last_error", "tpsc_find_mu0", "tpsc_chi0", "tpsc_sum_rule",
    "tpsc_find_Usp_Uch_grid", "tpsc_find_Usp_Uch", "tpsc_vertex_rt", "tpsc_self_energy",
    "tpsc_calc_self_energy_full", "tpsc_trace_self_G", "tpsc_trace_diff_G", "tpsc_find_mu",
    "tpsc_chijj", "tpsc_solve", "tpsc_get", "tpsc_field_size", "tpsc_get_scalar",
    "tpsc_prof_enable", "tpsc_prof_reset", "tpsc_prof_count", "tpsc_prof_get",
    "tpsc_keep_chi_tau", "tpsc_device_bytes", "tpsc_pade_coef", "tpsc_pade_eval", "tpsc_pade_conductivity",
    "tpsc_get_params", "tpsc_write_files", "tpsc_gwn_local", "tpsc_fix_mu0", "tpsc_sigma_loss_forward",
    "tpsc_sigma_loss_backward", "tpsc_loss_forward", "tpsc_loss_backward", "tpsc_random_weights", "tpsc_sp_corr_length", "tpsc_solve_batch",
    "tpsc_batch_tune", "tpsc_batch_release", "tpsc_batch_points", "tpsc_set_params", "tpsc_find_zero", "tpsc_sum_rule_noalloc",
    "tpsc_mu0_infinite", "tpsc_fermi_surface", "tpsc_fs_lattice", "tpsc_mdc", "tpsc_edc", "tpsc_self_real_freq",
    "tpsc_save_self_bin", "tpsc_save_chi0_bin", "tpsc_read_self_bin", "tpsc_read_chi0_bin",
]

# double f(double x, void* user) of tpsc_find_zero
ROOT_FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)


def lib():
    """Load libtpsc_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TPSCError(f"{LIB_PATH} not built: run __graft_entry__.build() (make -C tpsc_conductivity_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.tpsc_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.tpsc_destroy.argtypes = [vp]
    L.tpsc_destroy.restype = None
    L.tpsc_last_error.argtypes = [vp]
    L.tpsc_last_error.restype = C.c_char_p
    L.tpsc_find_mu0.argtypes = [vp, dp]
    L.tpsc_chi0.argtypes = [vp, C.c_int, C.c_int]
    L.tpsc_sum_rule.argtypes = [vp, C.c_int, C.c_double, dp]
    L.tpsc_find_Usp_Uch_grid.argtypes = [vp, C.c_int, C.c_int, dp, dp]
    L.tpsc_find_Usp_Uch.argtypes = [vp, dp, dp]
    L.tpsc_vertex_rt.argtypes = [vp]
    L.tpsc_self_energy.argtypes = [vp]
    L.tpsc_calc_self_energy_full.argtypes = [vp]
    L.tpsc_trace_self_G.argtypes = [vp, dp, dp]
    L.tpsc_trace_diff_G.argtypes = [vp, C.c_double, dp]
    L.tpsc_find_mu.argtypes = [vp, dp]
    L.tpsc_chijj.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.tpsc_solve.argtypes = [vp]
    L.tpsc_get.argtypes = [vp, C.c_int, vp, C.c_size_t]
    L.tpsc_field_size.argtypes = [vp, C.c_int]
    L.tpsc_field_size.restype = C.c_size_t
    L.tpsc_get_scalar.argtypes = [vp, C.c_int, dp]
    L.tpsc_prof_enable.argtypes = [vp, C.c_int]
    L.tpsc_prof_reset.argtypes = [vp]
    L.tpsc_prof_count.argtypes = [vp]
    L.tpsc_prof_get.argtypes = [vp, C.c_int, C.POINTER(C.c_char_p), dp, ip]
    L.tpsc_keep_chi_tau.argtypes = [vp, C.c_int]
    L.tpsc_device_bytes.argtypes = [vp, dp]
    L.tpsc_pade_coef.argtypes = [dp, dp, C.c_int, dp]
    L.tpsc_pade_eval.argtypes = [dp, dp, C.c_int, dp, C.c_int, dp]
    L.tpsc_pade_conductivity.argtypes = [dp, C.c_int, C.c_double, C.c_int, dp, C.c_int, C.c_double, dp]
    L.tpsc_get_params.argtypes = [vp, C.POINTER(Params)]
    L.tpsc_write_files.argtypes = [vp, C.c_char_p, ip]
    L.tpsc_gwn_local.argtypes = [vp, dp]
    L.tpsc_fix_mu0.argtypes = [vp, C.c_double]
    L.tpsc_sigma_loss_forward.argtypes = [vp, dp, dp]
    L.tpsc_sigma_loss_backward.argtypes = [vp, C.c_double, dp]
    L.tpsc_loss_forward.argtypes = [vp, dp, C.c_int, dp]
    L.tpsc_loss_backward.argtypes = [vp, C.c_double, dp]
    L.tpsc_random_weights.argtypes = [C.c_ulonglong, C.c_size_t, dp]
    L.tpsc_sp_corr_length.argtypes = [vp, dp]
    L.tpsc_solve_batch.argtypes = [C.POINTER(Params), C.c_int, C.c_int, C.POINTER(BatchResult), dp, C.c_int]
    L.tpsc_batch_tune.argtypes = [C.POINTER(Params), C.c_int, ip, dp]
    L.tpsc_batch_release.argtypes = []
    L.tpsc_batch_points.argtypes = [C.c_ulonglong, C.c_int, C.POINTER(Params)]
    L.tpsc_set_params.argtypes = [vp, C.POINTER(Params)]
    L.tpsc_find_zero.argtypes = [ROOT_FN, vp, dp, dp, dp, ip]
    L.tpsc_mu0_infinite.argtypes = [vp, dp]
    L.tpsc_fermi_surface.argtypes = [vp, C.c_int, dp, ip]
    L.tpsc_fs_lattice.argtypes = [vp, C.c_int, ip, dp]
    L.tpsc_mdc.argtypes = [vp, ip, C.c_int, C.c_double, dp, C.c_int, C.c_int, dp, dp]
    L.tpsc_edc.argtypes = [vp, ip, ip, C.c_int, dp, C.c_int, C.c_int, C.c_int, dp, C.c_int, dp]
    L.tpsc_self_real_freq.argtypes = [vp, ip, ip, C.c_int, dp, C.c_int, dp, C.c_int, dp, dp, dp, dp]
    L.tpsc_save_self_bin.argtypes = [vp, C.c_char_p]
    L.tpsc_save_chi0_bin.argtypes = [vp, C.c_char_p]
    L.tpsc_read_self_bin.argtypes = [C.c_char_p, C.c_int, C.c_int, dp]
    L.tpsc_read_chi0_bin.argtypes = [C.c_char_p, ip, ip, dp, C.c_size_t]
    # device addresses (chi0(q,iqn), its tails, the value) and the stream as plain pointers
    L.tpsc_sum_rule_noalloc.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_int, C.c_double, vp, vp]
    _lib = L
    return L


def exported_symbols():
    """Names of include/tpsc_capi.h entry points present in the loaded library."""
    L = lib()
    return [n for n in EXPORTS if hasattr(L, n)]


class Context:
    """Owns one tpsc_ctx (device buffers + stream for one parameter point)."""

    def __init__(self, nk0, Nw0, n, T, U, tp=0.0, tpp=0.0, t3=0.0, eta=0.1, coarse_pass=True, precision=0, t=1.0):
        self._h = C.c_void_p(None)
        self.params = Params(nk0, Nw0, n, T, eta, tp, tpp, t3, U, int(bool(coarse_pass)), precision, t)
        rc = lib().tpsc_create(C.byref(self.params), C.byref(self._h))
        if rc != TPSC_OK:
            msg = self.last_error()
            self.close()
            raise TPSCError(f"tpsc_create failed ({ERRORS.get(rc, rc)}): {msg}")

    def last_error(self):
        if not self._h:
            return "no context"
        m = lib().tpsc_last_error(self._h)
        return m.decode() if m else ""

    def check(self, rc, what):
        if rc != TPSC_OK:
            raise TPSCError(f"{what} failed ({ERRORS.get(rc, rc)}): {self.last_error()}")

    def close(self):
        if self._h:
            lib().tpsc_destroy(self._h)
            self._h = C.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def scalar(self, name):
        v = C.c_double(0.0)
        self.check(lib().tpsc_get_scalar(self._h, SCALARS[name], C.byref(v)), f"tpsc_get_scalar({name})")
        return v.value

    def write_files(self, directory=None):
        """The reference's text outputs for the completed stages (tpsc_write_files); returns the count."""
        n = C.c_int(0)
        d = directory.encode() if directory else None
        self.check(lib().tpsc_write_files(self._h, d, C.byref(n)), "tpsc_write_files")
        return n.value

    def get(self, name):
        f = FIELDS[name]
        nbytes = lib().tpsc_field_size(self._h, f)
        out = np.empty(nbytes // 8, dtype=np.float64)
        self.check(lib().tpsc_get(self._h, f, out.ctypes.data_as(C.c_void_p), nbytes), f"tpsc_get({name})")
        if name in ("Self", "Sigma_kw"):
            return out.view(np.complex128)
        return out


# ---- analytic continuation (host FP64, no device needed) ---------------------------------
def _cplx(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def pade_coef(z, f):
    """Thiele continued-fraction coefficients through (z_i, f_i) (generique.h:112)."""
    z, zp = _cplx(z)
    f, fp = _cplx(f)
    coef = np.zeros(len(z), dtype=np.complex128)
    rc = lib().tpsc_pade_coef(zp, fp, len(z), coef.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise TPSCError(f"tpsc_pade_coef failed ({ERRORS.get(rc, rc)})")
    return coef


def pade_eval(z, coef, x):
    """The approximant of pade_coef at the points x (generique.cpp:1630-1648)."""
    z, zp = _cplx(z)
    coef, cp = _cplx(coef)
    x, xp = _cplx(np.atleast_1d(x))
    out = np.zeros(len(x), dtype=np.complex128)
    rc = lib().tpsc_pade_eval(zp, cp, len(z), xp, len(x), out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise TPSCError(f"tpsc_pade_eval failed ({ERRORS.get(rc, rc)})")
    return out


def pade_conductivity(lam, T, npade, omega, eta):
    """sigma(w) = Im Lambda(w + i eta)/w, Lambda continued from Lambda(iq_n) = lam[n], n < npade."""
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    omega = np.ascontiguousarray(np.atleast_1d(omega), dtype=np.float64)
    sig = np.zeros(len(omega))
    dpp = C.POINTER(C.c_double)
    rc = lib().tpsc_pade_conductivity(lam.ctypes.data_as(dpp), len(lam), T, npade, omega.ctypes.data_as(dpp),
                                      len(omega), eta, sig.ctypes.data_as(dpp))
    if rc:
        raise TPSCError(f"tpsc_pade_conductivity failed ({ERRORS.get(rc, rc)})")
    return sig


def find_zero(f, init, lims=None):
    """tpsc_find_zero (host root search, no device): (found, root, evaluations)."""
    cb = ROOT_FN(lambda x, _u: float(f(x)))
    dpp = C.POINTER(C.c_double)
    ini = np.ascontiguousarray(init, dtype=np.float64)
    lim = np.ascontiguousarray(lims, dtype=np.float64) if lims is not None else None
    root, ne = C.c_double(0.0), C.c_int(0)
    rc = lib().tpsc_find_zero(cb, None, ini.ctypes.data_as(dpp), lim.ctypes.data_as(dpp) if lim is not None else None,
                              C.byref(root), C.byref(ne))
    if rc not in (TPSC_OK, -5):
        raise TPSCError(f"tpsc_find_zero failed ({ERRORS.get(rc, rc)})")
    return rc == TPSC_OK, root.value, ne.value


def random_weights(seed, n):
    """n weights ~ U(0,1) from std::mt19937_64(seed) (tpsc_random_weights)."""
    out = np.zeros(n)
    rc = lib().tpsc_random_weights(seed, n, out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise TPSCError(f"tpsc_random_weights failed ({ERRORS.get(rc, rc)})")
    return out


def batch_points(seed, n):
    """The random parameter points of tpsc_batch_points (a ctypes Params array)."""
    pts = (Params * n)()
    rc = lib().tpsc_batch_points(seed, n, pts)
    if rc:
        raise TPSCError(f"tpsc_batch_points failed ({ERRORS.get(rc, rc)})")
    return pts


def solve_batch(pts, workers=0, ldc=0):
    """tpsc_solve_batch over a Params array: (results array, Lambda_jj rows or None)."""
    n = len(pts)
    res = (BatchResult * n)()
    jj = np.zeros((n, ldc)) if ldc else None
    jp = jj.ctypes.data_as(C.POINTER(C.c_double)) if ldc else None
    rc = lib().tpsc_solve_batch(pts, n, workers, res, jp, ldc)
    if rc:
        raise TPSCError(f"tpsc_solve_batch failed ({ERRORS.get(rc, rc)})")
    return res, jj


def batch_release():
    """tpsc_batch_release: frees the contexts the queued batch path keeps between calls."""
    lib().tpsc_batch_release()


def batch_tune(pts):
    w = C.c_int(0)
    rates = np.zeros(4)
    rc = lib().tpsc_batch_tune(pts, len(pts), C.byref(w), rates.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise TPSCError(f"tpsc_batch_tune failed ({ERRORS.get(rc, rc)})")
    return w.value, rates


def load_Self_bin(path, N, Nw):
    """green::load_Self_bin (green.cpp:1516-1540): the (Nw/2+1) x nk8 complex array of a self_bin_* file."""
    nk = N // 2 + 1
    nk8 = nk * (nk + 1) // 2
    out = np.zeros((Nw // 2 + 1) * nk8, dtype=np.complex128)
    rc = lib().tpsc_read_self_bin(str(path).encode(), N, Nw, out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise TPSCError(f"tpsc_read_self_bin failed ({ERRORS.get(rc, rc)}): {path}")
    return out.reshape(Nw // 2 + 1, nk8)


def loadChi0(path):
    """chi0::loadChi0 (chi0.cpp:99-156): (nbq, nwmax, chi0 array (nwmax+1, nbq(nbq+1)/2)) of a chi0_* file."""
    nbq, nwmax = C.c_int(0), C.c_int(0)
    L = lib()
    rc = L.tpsc_read_chi0_bin(str(path).encode(), C.byref(nbq), C.byref(nwmax), None, 0)
    if rc:
        raise TPSCError(f"tpsc_read_chi0_bin failed ({ERRORS.get(rc, rc)}): {path}")
    nk8 = nbq.value * (nbq.value + 1) // 2
    out = np.zeros((nwmax.value + 1) * nk8)
    rc = L.tpsc_read_chi0_bin(str(path).encode(), C.byref(nbq), C.byref(nwmax), out.ctypes.data_as(C.POINTER(C.c_double)),
                              out.size)
    if rc:
        raise TPSCError(f"tpsc_read_chi0_bin failed ({ERRORS.get(rc, rc)}): {path}")
    return nbq.value, nwmax.value, out.reshape(nwmax.value + 1, nk8)
