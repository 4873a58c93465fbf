"""Algorithmic HBM bytes of the TPSC kernels.

Bytes count compulsory reads and writes of the HBM-resident arrays each kernel family must
touch (LDS and register intermediates count zero; per-slice staging the kernel adds for
itself does not count).  Sizes: nk = N/2+1, nk8 = nk(nk+1)/2 (wedge), nwb = Nw/2+1,
nt = Nw+1 (tau points), nq = nk(nk-2) (Lambda grid).  Element sizes by path:
s = 8 (FP64 path) or 4 (mixed path) for the tau arrays chi0(q,tau), chi0(r,tau), V,
Sigma(k,tau), h, f0; sc = 16 or 8 for the stored Sigma(k,iwn); chi0(q,iqn) is FP64 (8) in
both paths.
"""

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)

SLAB_KERNELS = ("chi0_slab", "vertex_slab", "sigma_slab", "jj_hf0_slab", "jj_hVf0V_slab")


def sizes(N, Nw):
    nk = N // 2 + 1
    return dict(nk=nk, nk8=nk * (nk + 1) // 2, nwb=Nw // 2 + 1, nt=Nw + 1, nq=nk * (nk - 2), Nw=Nw)


def elem_sizes(precision):
    """(s, sc): bytes of one tau-array element and of one stored Sigma(k,iwn) element"""
    return (4, 8) if precision == "mixed" else (8, 16)


def kernel_bytes(N, Nw, precision="fp64"):
    """bytes per launch of each kernel family"""
    z = sizes(N, Nw)
    nk8, nwb, nt, nq = z["nk8"], z["nwb"], z["nt"], z["nq"]
    s, sc = elem_sizes(precision)
    d = 8
    return {
        # write chi0(q,tau) and chi0(r,tau) for slices 1..nwb-1
        "chi0_slab": 2 * (nwb - 1) * nk8 * s,
        # read chi0(q,tau), write chi0(q,iqn) (FP64) and the two tail arrays
        "chi0_pencil": nwb * nk8 * s + nwb * nk8 * d + 2 * nk8 * d,
        # read chi0(q,iqn), write V(r,iqn)
        "vertex_slab": nwb * nk8 * d + nwb * nk8 * s,
        # read V(r,iqn) and chi0(r,tau), write V(r,tau)
        "vertex_pencil": 3 * nwb * nk8 * s,
        # read V(r,|tau|) per slice, write Sigma(k,tau)
        "sigma_slab": 2 * nt * nk8 * s,
        # read Sigma(k,tau), write Sigma(k,iwn) and three moment arrays
        "sigma_pencil": nt * nk8 * s + (Nw // 2) * nk8 * sc + 3 * nk8 * d,
        # write h(r,tau), f0(r,tau)
        "jj_hf0_slab": 2 * nt * nq * s,
        # read and write h, f0; read V(r,|tau|)
        "jj_hVf0V_slab": 4 * nt * nq * s + nt * nk8 * s,
        # read hV, f0V and (once) Sigma(k,iwn); partial rows out
        "jj_pencil": 2 * nt * nq * s + (Nw // 2) * nk8 * sc + 2 * ((nq + 15) // 16) * 2 * nwb * d,
        # one sum-rule evaluation: read chi0(q,iqn)
        "sum_rule_eval": nwb * nk8 * d,
        # one Tr(G-G0) evaluation: read Sigma(k,iwn)
        "mu_eval": (Nw // 2) * nk8 * sc,
    }


def solve_bytes(N, Nw, n_sum_rule_evals, n_mu_evals, precision="fp64"):
    """algorithmic bytes of one solve: chi0 -> roots -> V -> Sigma -> mu -> Lambda_jj"""
    kb = kernel_bytes(N, Nw, precision)
    fixed = sum(kb[k] for k in ("chi0_slab", "chi0_pencil", "vertex_slab", "vertex_pencil", "sigma_slab",
                                "sigma_pencil", "jj_hf0_slab", "jj_hVf0V_slab", "jj_pencil"))
    return fixed + n_sum_rule_evals * kb["sum_rule_eval"] + n_mu_evals * kb["mu_eval"]
