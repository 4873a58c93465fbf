"""CPU checks of the drop-in boundary: libtpsc_hip.so loads, exports every entry point that
include/tpsc_capi.h declares, and rejects bad input before touching a device."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "tpsc_capi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tpsc_[a-zA-Z0-9_]+)\s*\(", txt)))


def test_header_declares_the_stage_surface():
    names = header_functions()
    for n in ("tpsc_create", "tpsc_destroy", "tpsc_find_mu0", "tpsc_chi0", "tpsc_sum_rule", "tpsc_find_Usp_Uch",
              "tpsc_vertex_rt", "tpsc_self_energy", "tpsc_trace_self_G", "tpsc_find_mu", "tpsc_chijj", "tpsc_get"):
        assert n in names


def test_library_exports_every_declared_symbol(hip_lib):
    missing = [n for n in header_functions() if not hasattr(hip_lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from tpsc_conductivity_amd import capi

    assert set(header_functions()) <= set(capi.EXPORTS)


def _params(**kw):
    from tpsc_conductivity_amd.capi import Params

    d = dict(nk0=16, Nw0=64, n=0.9, T=0.1, eta=0.1, tp=0.0, tpp=0.0, t3=0.0, U=2.0, coarse_pass=0, precision=0)
    d.update(kw)
    return Params(**d)


@pytest.mark.parametrize("bad", [dict(nk0=20), dict(nk0=512), dict(Nw0=100), dict(precision=2), dict(T=0.0),
                                 dict(coarse_pass=1, nk0=4), dict(n=0.0), dict(n=2.0), dict(n=2.5),
                                 dict(n=-0.1)])
def test_invalid_parameters_are_rejected_without_a_device(hip_lib, bad):
    h = C.c_void_p(None)
    p = _params(**bad)
    rc = hip_lib.tpsc_create(C.byref(p), C.byref(h))
    assert rc == -1  # TPSC_E_INVALID
    msg = hip_lib.tpsc_last_error(h)
    assert msg
    hip_lib.tpsc_destroy(h)


def test_null_arguments(hip_lib):
    assert hip_lib.tpsc_create(None, None) == -1
    assert hip_lib.tpsc_chi0(None, 16, 64) == -1
    assert hip_lib.tpsc_chijj(None, 0, 3, 1) == -1


def test_noalloc_sum_rule_rejects_bad_sizes_without_a_device(hip_lib):
    """tpsc_sum_rule_noalloc checks its sizes before any device call: N beyond its fixed
    scratch array, unsupported grids and a bad channel give TPSC_E_INVALID and a message in
    tpsc_last_error(NULL)."""
    x = C.c_void_p(16)  # never dereferenced
    f = hip_lib.tpsc_sum_rule_noalloc
    assert f(x, x, x, 512, 256, 0.1, 0.9, 4.0, 0.0, 0, 1.0, x, None) == -1
    assert b"N=512" in hip_lib.tpsc_last_error(None)
    assert f(x, x, x, 20, 256, 0.1, 0.9, 4.0, 0.0, 0, 1.0, x, None) == -1
    assert f(x, x, x, 64, 100, 0.1, 0.9, 4.0, 0.0, 0, 1.0, x, None) == -1
    assert f(x, x, x, 64, 256, 0.1, 0.9, 4.0, 0.0, 2, 1.0, x, None) == -1
    assert b"channel" in hip_lib.tpsc_last_error(None)
    assert f(None, x, x, 64, 256, 0.1, 0.9, 4.0, 0.0, 0, 1.0, x, None) == -1
