#!/usr/bin/env python3
"""Headline benchmark: TPSC solves/s at 256^2 x 2048 iwn on one MI355X, mixed precision
against the FP64 GPU path.

One step = one full TPSC solve on the GPU: mu0 -> chi0(q,iqn) -> Usp/Uch sum-rule roots ->
V(r,tau) -> Sigma(k,iwn) -> interacting mu -> Lambda_jj(iqn) bubble + Maki-Thompson + kx0
(the coarse diagnostic pass of chi.cpp:285-305 is not part of a step).  Inputs are the
model parameters; every array lives in HBM for the whole solve.

`value` is the mixed-precision path (FP32 slab transforms and tau arrays; FP64 boundary
data, tail moments, roots and Matsubara sums).  The FP64 path is timed in the same run, and
every output of the mixed path is compared with it (max |mixed - fp64| / max |fp64|).

Prints one JSON line.  Per-kernel times come from HIP events the library records on its
own streams during one extra profiled solve after the timed region.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TPSC Σ/χ/Λjj solves/s at 256²×2048 iωn: mixed vs FP64 GPU at 1e-6; HBM GB/s"
WORKLOAD = dict(N=256, Nw=2048, n=0.9, T=0.05, U=5.0, tp=-0.2, tpp=0.1, t3=0.0)
TOL = 1e-6
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")
FLOPS_FILE = os.path.join(ROOT, "profiles", "flops.json")


def cpu_baseline(cfg, repeats=3):
    """The same workload in the FP64 C oracle on the host cores: built -O3 -march=native on this
    host, one warm-up solve, then the median of `repeats` timed solves, with the median time of
    each stage (A mu0 + chi0, B Usp/Uch roots, C vertex, D self-energy, E traceSelfG + mu,
    F Lambda_jj)."""
    import statistics

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    build = "-O3 -march=native"
    try:
        oracle.use_native()
    except Exception as e:  # no compiler on this host: the portable x86-64-v3 build
        build = f"-O3 -march=x86-64-v3 (native build failed: {e})"
    threads = oracle.omp_max_threads()
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    o = oracle.Oracle(cfg["N"], cfg["Nw"], cfg["n"], cfg["T"], cfg["U"], tp=cfg["tp"], tpp=cfg["tpp"],
                      t3=cfg["t3"], coarse_pass=False)
    o.solve_timed()
    print("[bench] cpu baseline warm-up done", file=sys.stderr, flush=True)
    totals, stages = [], []
    for r in range(repeats):
        t0 = time.perf_counter()
        ok, st = o.solve_timed()
        totals.append(time.perf_counter() - t0)
        stages.append(st)
        print(f"[bench] cpu baseline solve {r + 1}/{repeats}: {totals[-1]:.2f} s", file=sys.stderr, flush=True)
    med = statistics.median(totals)
    names = ("A_mu0_chi0", "B_Usp_Uch", "C_vertex", "D_self_energy", "E_traces_mu", "F_Lambda_jj")
    return {"value": 1.0 / med, "unit": "solves/s", "cores": threads, "kind": "port",
            "nproc": nproc, "omp_max_threads": threads, "build": build,
            "stages_s": {nm: round(statistics.median(s[i] for s in stages), 4) for i, nm in enumerate(names)},
            "solves_s": [round(t, 3) for t in totals],
            "sample": f"median of {repeats} full {cfg['N']}^2 x {cfg['Nw']} solves (mu0..Lambda_jj) of the FP64 C "
                      f"oracle ({build}) after one warm-up solve, OpenMP {threads} threads on {nproc} host cores, "
                      f"{med:.2f} s per solve"}


def make(cfg, precision):
    from tpsc_conductivity_amd.tpsc import CondOpt

    return CondOpt(cfg["N"], cfg["Nw"], cfg["n"], cfg["T"], params=(cfg["tp"], cfg["tpp"], cfg["t3"], cfg["U"]),
                   coarse_pass=False, precision=precision)


def timed(g, steps, warmup, tag):
    for i in range(warmup):
        g.solve()
        print(f"[bench] {tag} warmup {i + 1}/{warmup} done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        g.solve()  # returns after the stream is synchronised (scalars copied to the host)
    ms = 1e3 * (time.perf_counter() - t0) / steps
    print(f"[bench] {tag}: timed {steps} solves: {ms:.3f} ms/solve", file=sys.stderr, flush=True)
    return ms


def outputs(g):
    import numpy as np

    chi0 = g.get("chi0")
    Usp, Uch = g.scalar("Usp"), g.scalar("Uch")
    return {"chi0": chi0, "chisp": chi0 / (1 - Usp * chi0 / 2), "chich": chi0 / (1 + Uch * chi0 / 2),
            "Sigma_kw": g.get("Sigma_kw"), "Lambda_jj": g.get("chijj1") + g.get("chijj2"),
            "Usp": np.array([Usp]), "Uch": np.array([Uch]), "mu": np.array([g.scalar("mu")]),
            "kx0": np.array([g.scalar("kx0")])}


def rel_errors(a, b):
    import numpy as np

    return {k: float(np.max(np.abs(a[k] - b[k])) / np.max(np.abs(b[k]))) for k in b}


def profile(g, N, Nw, precision):
    from tpsc_conductivity_amd import roofline

    g.profile(True)
    g.profile_reset()
    g.solve()
    table = g.profile_table()
    g.profile(False)
    kb = roofline.kernel_bytes(N, Nw, precision)
    per = {}
    for name, (tot_ms, calls) in table.items():
        e = {"avg_ms": tot_ms / calls, "calls": calls}
        if name in kb:
            e["GBps"] = kb[name] / (e["avg_ms"] * 1e-3) / 1e9
        per[name] = e
    return table, per, kb


def bench_al(cfg, args):
    """Lambda_jj including the Aslamazov-Larkin term (chi3) after one solve: wall time of the
    call and, from the library's HIP events, the AL set-up and per-frequency times."""
    g = make(cfg, args.precision)
    g.solve()
    for i in range(max(args.warmup, 1)):
        g.calc_chijj_vertex_corr_all_optim(chi3=True, mmin=args.al)
        print(f"[bench] al warmup {i + 1} done", file=sys.stderr, flush=True)
    steps = max(1, min(args.steps, 3))
    t0 = time.perf_counter()
    for _ in range(steps):
        g.calc_chijj_vertex_corr_all_optim(chi3=True, mmin=args.al)
    ms = 1e3 * (time.perf_counter() - t0) / steps
    g.profile(True)
    g.profile_reset()
    g.calc_chijj_vertex_corr_all_optim(chi3=True, mmin=args.al)
    table = g.profile_table()
    g.profile(False)
    nf = int(g.scalar("al_N")) - 1
    al_ms = table["al_setup"][0] + table["al_freq"][0]
    print(json.dumps({"metric": f"Aslamazov-Larkin Lambda_jj3 boson frequencies/s at {cfg['N']}²×{cfg['Nw']} iωn",
                      "value": nf / (al_ms * 1e-3), "unit": "frequencies/s", "higher_is_better": True,
                      "n_freq": nf, "mmin": args.al, "ms_per_freq": table["al_freq"][0] / max(nf, 1),
                      "al_setup_ms": table["al_setup"][0], "al_ms": al_ms, "chijj_call_ms": ms,
                      "steps": steps, "dtype": "f64", "precision": args.precision, "data": "synthetic",
                      "config": {"workload": f"TPSC {cfg['N']}x{cfg['N']} lattice, {cfg['Nw']} Matsubara "
                                             f"frequencies, n={cfg['n']}, T={cfg['T']}, U={cfg['U']}"}}),
          flush=True)


def bench_batch(args):
    """Throughput of tpsc_solve_batch on the 64 random points (N in {32..128}, T, n, U, t'), in
    mixed precision (value) and FP64: the tuned worker count, the queued path (one worker, no
    host wait inside a point) and the threaded path (the fastest of 2, 4, 8 workers)."""
    from tpsc_conductivity_amd import capi

    pts = capi.batch_points(64064, 64)
    paths = {}
    for prec, code in (("mixed", capi.PREC_MIXED), ("fp64", capi.PREC_FP64)):
        for p in pts:
            p.precision = code
        w, rates = capi.batch_tune(pts)
        wt = (2, 4, 8)[max(range(3), key=lambda k: rates[k + 1])]
        capi.solve_batch(pts, workers=w)  # warm-up (code objects, allocator)
        runs = {}
        for tag, nw in (("tuned", w), ("queued", 1), ("threaded", wt)):
            t0 = time.perf_counter()
            res, _ = capi.solve_batch(pts, workers=nw)
            dt = time.perf_counter() - t0
            st = [r.status for r in res]
            runs[tag] = {"workers": nw, "seconds": dt, "solved": st.count(0),
                         "unsupported": st.count(-1), "failed": len(st) - st.count(0) - st.count(-1),
                         "finished_on_queue": sum(r.queued for r in res)}
            print(f"[bench] batch {prec} {tag}: {nw} workers, {dt:.3f} s", file=sys.stderr, flush=True)
        paths[prec] = {"solves_per_s": runs["tuned"]["solved"] / runs["tuned"]["seconds"],
                       "queued_over_threaded": runs["threaded"]["seconds"] / runs["queued"]["seconds"],
                       "tuner_points_per_s": {str(k): float(v) for k, v in zip((1, 2, 4, 8), rates)},
                       "runs": runs}
    print(json.dumps({"metric": "TPSC batched solves/s, 64 random (N, T, n, U, t') points, one MI355X",
                      "value": paths["mixed"]["solves_per_s"], "unit": "solves/s", "higher_is_better": True,
                      "mixed": paths["mixed"], "fp64": paths["fp64"], "dtype": "f32+f64", "data": "synthetic",
                      "config": {"workload": "tpsc_batch_points(64064, 64): N in {32,48,64,96,128}, "
                                             "T in [0.05,0.5], n in [0.8,1], U in [1,6], t' in [-0.3,0], "
                                             "Nw = smallest power of two with (Nw-1) pi T >= 20"}}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--N", type=int, default=WORKLOAD["N"])
    ap.add_argument("--Nw", type=int, default=WORKLOAD["Nw"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--kernel-table", action="store_true", help="print per-kernel times to stderr")
    ap.add_argument("--precision", choices=("mixed", "fp64"), default="mixed",
                    help="path timed for `value` (default: mixed)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run an elementwise copy of a 1 GiB device buffer 5 times (reference kernel "
                         "for the FETCH_SIZE / WRITE_SIZE scale of the traffic figure) and exit")
    ap.add_argument("--al", type=int, default=None, metavar="MMIN",
                    help="instead of the headline line, time the Aslamazov-Larkin term on the sparse boson set "
                         "of parameter MMIN (one solve first, then Lambda_jj with chi3) and print its own line")
    ap.add_argument("--batch", action="store_true",
                    help="instead of the headline line, solve the 64 random points of tpsc_batch_points(64064) "
                         "with the tuned worker count and with one worker, and print their own line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(a parameter point does not span GPUs here)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        import torch

        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8,
                          "launches": 5}), flush=True)
        return

    if args.batch:
        return bench_batch(args)
    if args.al is not None:
        return bench_al(dict(WORKLOAD, N=args.N, Nw=args.Nw), args)

    from tpsc_conductivity_amd import roofline

    cfg = dict(WORKLOAD, N=args.N, Nw=args.Nw)
    prec = args.precision
    g = make(cfg, prec)
    ms = timed(g, args.steps, args.warmup, prec)
    table, per, kb = profile(g, cfg["N"], cfg["Nw"], prec)
    n_sr = int(g.scalar("nevals_sp") + g.scalar("nevals_ch"))
    n_mu = int(g.scalar("nevals_mu"))
    out_main = outputs(g)
    g.close()
    if args.kernel_table:
        for name, v in sorted(table.items(), key=lambda x: -x[1][0]):
            extra = f"  {per[name]['GBps']:8.1f} GB/s" if "GBps" in per[name] else ""
            print(f"[bench] {prec} {name:16s} {v[0]:9.3f} ms  x{v[1]}{extra}", file=sys.stderr)
    dominant = max((k for k in per if k in kb), key=lambda k: table[k][0])
    d = per[dominant]
    slab_ms = sum(table[k][0] for k in roofline.SLAB_KERNELS if k in table)
    slab_b = sum(kb[k] * table[k][1] for k in roofline.SLAB_KERNELS if k in table)
    key = f"{prec}:{dominant}:{cfg['N']}x{cfg['Nw']}"
    traffic, valu = None, None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        if key in tr:
            traffic = tr[key]["bytes_per_launch"]
    if os.path.exists(FLOPS_FILE):
        with open(FLOPS_FILE) as f:
            fl = json.load(f)
        if key in fl:
            # FP64 work of the same kernel: its counted VALU FP64 instructions (per launch, a
            # fixed figure for this workload) over the launch time measured here
            f64 = fl[key]["fp64_flops_per_launch"]
            tf = f64 / (d["avg_ms"] * 1e-3) / 1e12
            valu = {"bound": "fp64 valu", "kernel": dominant, "achieved": tf, "peak": roofline.FP64_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": tf / roofline.FP64_PEAK_TFLOPS, "fp64_flops_per_launch": f64,
                    "flops_source": "64 x (ADD + MUL + TRANS + 2 FMA) of SQ_INSTS_VALU_*_F64 per launch, "
                                    "profiles/flops.json"}
    out = {
        "metric": METRIC,
        "value": 1e3 / ms,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if prec == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"TPSC solve N={cfg['N']} Nw={cfg['Nw']} n={cfg['n']} T={cfg['T']} U={cfg['U']} "
                               f"t'={cfg['tp']} t''={cfg['tpp']}, {prec} precision, mu0..Lambda_jj bubble+MT, "
                               "no AL",
                   "N": cfg["N"], "Nw": cfg["Nw"], "precision": prec},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": d["GBps"], "peak": roofline.HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": d["GBps"] / roofline.HBM_PEAK_GBPS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": kb[dominant], "avg_launch_ms": d["avg_ms"]},
        "valu_roofline": valu,
        "slab_passes": {"kernels": list(roofline.SLAB_KERNELS), "ms": slab_ms,
                        "GBps": slab_b / (slab_ms * 1e-3) / 1e9 if slab_ms > 0 else None,
                        "frac": slab_b / (slab_ms * 1e-3) / 1e9 / roofline.HBM_PEAK_GBPS if slab_ms > 0 else None},
        "solve_hbm_GBps": roofline.solve_bytes(cfg["N"], cfg["Nw"], n_sr, n_mu, prec) / (ms * 1e-3) / 1e9,
        "kernels_ms": {k: round(v["avg_ms"] * v["calls"], 4) for k, v in per.items()},
    }
    if prec == "mixed" and not args.no_fp64:
        g = make(cfg, "fp64")
        ms64 = timed(g, args.steps, args.warmup, "fp64")
        out_64 = outputs(g)
        g.close()
        errs = rel_errors(out_main, out_64)
        out["fp64"] = {"ms_per_step": ms64, "solves_per_s": 1e3 / ms64, "speedup_mixed_over_fp64": ms64 / ms}
        out["mixed_rel_err_vs_fp64"] = {k: float(f"{v:.3e}") for k, v in errs.items()}
        out["mixed_within_tol"] = bool(all(v <= TOL for v in errs.values()))
        out["tol"] = TOL
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
