#!/usr/bin/env python3
"""Headline benchmark: TPSC solves/s at 256^2 x 2048 iwn on one MI355X (FP64 path).

One step = one full TPSC solve on the GPU: mu0 -> chi0(q,iqn) -> Usp/Uch sum-rule roots ->
V(r,tau) -> Sigma(k,iwn) -> interacting mu -> Lambda_jj(iqn) bubble + Maki-Thompson + kx0
(the coarse diagnostic pass of chi.cpp:285-305 is not part of a step).  Inputs are the
model parameters; every array lives in HBM for the whole solve.

Prints one JSON line.  Per-kernel times come from HIP events the library records on its
own stream during one extra profiled solve after the timed region.
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


def cpu_baseline(cfg):
    """One full solve of the same workload in the FP64 C oracle on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = oracle.Oracle(cfg["N"], cfg["Nw"], cfg["n"], cfg["T"], cfg["U"], tp=cfg["tp"], tpp=cfg["tpp"],
                      t3=cfg["t3"], coarse_pass=False)
    t0 = time.perf_counter()
    o.solve()
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"one full {cfg['N']}^2 x {cfg['Nw']} solve (mu0..Lambda_jj) of the FP64 C oracle, "
                      f"OpenMP {threads} threads, {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--N", type=int, default=WORKLOAD["N"])
    ap.add_argument("--Nw", type=int, default=WORKLOAD["Nw"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel-table", action="store_true", help="print per-kernel times to stderr")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(a parameter point does not span GPUs here)", file=sys.stderr)
        sys.exit(2)

    from tpsc_conductivity_amd import roofline
    from tpsc_conductivity_amd.tpsc import CondOpt

    cfg = dict(WORKLOAD, N=args.N, Nw=args.Nw)
    g = CondOpt(cfg["N"], cfg["Nw"], cfg["n"], cfg["T"], params=(cfg["tp"], cfg["tpp"], cfg["t3"], cfg["U"]),
                coarse_pass=False)
    for i in range(args.warmup):
        g.solve()
        print(f"[bench] warmup {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.solve()  # returns after the stream is synchronised (scalars copied to the host)
    dt = time.perf_counter() - t0
    ms = 1e3 * dt / args.steps
    print(f"[bench] timed {args.steps} solves: {ms:.3f} ms/solve", file=sys.stderr, flush=True)

    # one profiled solve: HIP events around every kernel family on the library stream
    g.profile(True)
    g.profile_reset()
    g.solve()
    table = g.profile_table()
    g.profile(False)
    kb = roofline.kernel_bytes(cfg["N"], cfg["Nw"])
    per_kernel = {}
    for name, (tot_ms, calls) in table.items():
        if name in kb:
            avg = tot_ms / calls
            per_kernel[name] = {"avg_ms": avg, "calls": calls, "GBps": kb[name] / (avg * 1e-3) / 1e9}
        else:
            per_kernel[name] = {"avg_ms": tot_ms / calls, "calls": calls}
    if args.kernel_table:
        for name, v in sorted(table.items(), key=lambda x: -x[1][0]):
            extra = f"  {per_kernel[name].get('GBps', 0):8.1f} GB/s" if "GBps" in per_kernel[name] else ""
            print(f"[bench] {name:16s} {v[0]:9.3f} ms  x{v[1]}{extra}", file=sys.stderr)
    dominant = max((k for k in per_kernel if k in kb), key=lambda k: table[k][0])
    d = per_kernel[dominant]
    n_sr = int(g.scalar("nevals_sp") + g.scalar("nevals_ch"))
    n_mu = int(g.scalar("nevals_mu"))
    solve_b = roofline.solve_bytes(cfg["N"], cfg["Nw"], n_sr, n_mu)
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
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"TPSC solve N={cfg['N']} Nw={cfg['Nw']} n={cfg['n']} T={cfg['T']} U={cfg['U']} "
                               f"t'={cfg['tp']} t''={cfg['tpp']}, FP64, mu0..Lambda_jj bubble+MT, no AL",
                   "N": cfg["N"], "Nw": cfg["Nw"], "precision": "fp64"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": d["GBps"], "peak": roofline.HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": d["GBps"] / roofline.HBM_PEAK_GBPS, "traffic": None,
                     "algorithmic_bytes_per_launch": kb[dominant], "avg_launch_ms": d["avg_ms"]},
        "solve_hbm_GBps": solve_b / (ms * 1e-3) / 1e9,
        "kernels_ms": {k: round(v["avg_ms"] * v["calls"], 4) for k, v in per_kernel.items()},
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
