"""tpsc_sum_rule_noalloc (the allocation-free sum-rule evaluation over caller-owned device
arrays) against tpsc_sum_rule of a context holding the same chi0(q,iqn), through the C ABI: two
grids in a row, the second smaller, spin and charge channels at several vertices.  The
library's fixed scratch array keeps the first grid's weights and partials when the second call
starts, so the second grid checks that a call rewrites everything it reads.

The caller's device buffers come from the HIP runtime the library itself links
(/opt/rocm/lib/libamdhip64.so.7), called through ctypes: torch ships its own HIP runtime, which
cannot initialise the GPU once the library's runtime has (torch.cuda then sees no device)."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (N, Nw, n, T, U, t'): the larger grid first
CASES = [(64, 256, 0.9, 0.1, 4.0, -0.175), (16, 64, 1.05, 0.2, 3.0, -0.1)]
H2D, D2H = 1, 2


def hip_runtime():
    path = "/opt/rocm/lib/libamdhip64.so.7"
    if not os.path.exists(path):
        path = "libamdhip64.so"
    H = C.CDLL(path)
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return H


class DevBuf:
    def __init__(self, H, host):
        self.H, self.n = H, host.nbytes
        self.ptr = C.c_void_p()
        assert H.hipMalloc(C.byref(self.ptr), self.n) == 0
        assert H.hipMemcpy(self.ptr, host.ctypes.data, self.n, H2D) == 0

    def read(self, dtype=np.float64):
        out = np.empty(self.n // np.dtype(dtype).itemsize, dtype=dtype)
        assert self.H.hipMemcpy(out.ctypes.data, self.ptr, self.n, D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.ptr)


def test_noalloc_sum_rule_matches_context_sum_rule():
    from tpsc_conductivity_amd import capi

    L = capi.lib()
    H = hip_runtime()
    out = DevBuf(H, np.full(1, np.nan))
    for N, Nw, n, T, U, tp in CASES:
        print(f"[noalloc] N={N} Nw={Nw} n={n} T={T} U={U} tp={tp}", flush=True)
        ctx = capi.Context(N, Nw, n, T, U, tp=tp, coarse_pass=False)
        usp, uch = C.c_double(0.0), C.c_double(0.0)
        ctx.check(L.tpsc_find_Usp_Uch_grid(ctx.handle, N, Nw, C.byref(usp), C.byref(uch)), "find_Usp_Uch_grid")
        host = [np.ascontiguousarray(ctx.get(f)) for f in ("chi0", "chi0_inf", "chi0_inf2")]
        nk8 = (N // 2 + 1) * (N // 2 + 2) // 2
        assert host[0].size == (Nw // 2 + 1) * nk8 and host[1].size == nk8
        arrs = [DevBuf(H, a) for a in host]
        for channel, Ux in ((0, 0.5 * usp.value), (0, usp.value), (1, uch.value), (1, 2.0 * uch.value)):
            ref = C.c_double(0.0)
            ctx.check(L.tpsc_sum_rule(ctx.handle, channel, Ux, C.byref(ref)), "tpsc_sum_rule")
            assert H.hipMemcpy(out.ptr, np.full(1, np.nan).ctypes.data, 8, H2D) == 0
            rc = L.tpsc_sum_rule_noalloc(arrs[0].ptr, arrs[1].ptr, arrs[2].ptr, N, Nw, T, n, U, usp.value, channel, Ux,
                                         out.ptr, None)
            assert rc == 0, L.tpsc_last_error(None)
            got = float(out.read()[0])  # hipMemcpy on the null stream orders after the call's launches
            err = abs(got - ref.value) / max(1.0, abs(ref.value))
            print(f"[noalloc]   channel {channel} Ux={Ux:.6f}: {got!r} vs {ref.value!r} (rel {err:.1e})", flush=True)
            assert err <= 1e-12
        for a in arrs:
            a.free()
        ctx.close()
    out.free()
