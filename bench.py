"""Benchmark of the AMEn solve on one MI355X.

A step is one complete AMEn solve of the workload (all sweeps, to the solver's stopping
rule) with the operator and right-hand side already resident in HBM. The default workload
is the reference comparison case: d = 10, n = 64, A = Laplacian + potential (TT rank 4),
b = random rank-4 TT with ||b|| = 1, rank cap 80, eps 1e-8, FP64.

Prints one JSON line with the throughput (solves/s), p50/p99 latency, the worst relative
residual ||Ax - b|| / ||b|| of the timed solves, the roofline of the local operator (the
solver's dominant kernel work, timed with HIP events inside the library on its stream) and
the reference CPU solver's throughput on the same workload measured on this host.

    python bench.py [--steps K] [--warmup W] [--precision fp64|fp32] [--no-cpu-baseline]
"""
import argparse
import ctypes
import glob
import importlib.util
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/sec at rel. residual ≤1e-8, mixed vs FP64 plain; p50/p99 solve latency"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
HBM_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    p.add_argument("--d", type=int, default=10)
    p.add_argument("--n", type=int, default=64)
    p.add_argument("--rmax", type=int, default=80)
    p.add_argument("--eps", type=float, default=1e-8)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-threads", type=int, default=0)
    return p.parse_args()


def cpu_baseline(A, b, args):
    """The reference's own C++ AMEn (oracle/_ref, built from its sources) on this host's cores,
    one full solve of the same workload; the CPU oracle restatement if that build is absent."""
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    torch.set_num_threads(threads)
    d = len(A)
    so = glob.glob(os.path.join(ROOT, "oracle", "_ref", "torchttcpp*.so"))
    rA = [1] + [a.shape[3] for a in A]
    rb = [1] + [c.shape[2] for c in b]
    torch.manual_seed(0)
    if so:
        spec = importlib.util.spec_from_file_location("torchttcpp", so[0])
        ref = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(ref)
        t0 = time.perf_counter()
        ref.amen_solve(A, b, [], [a.shape[2] for a in A], rA, rb, [1] * (d + 1), 22, args.eps, args.rmax, 500,
                       4, 0, 40, 2, False, 0)
        dt = time.perf_counter() - t0
        kind = "reference"
    else:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import amen_oracle as O
        t0 = time.perf_counter()
        O.amen_solve([a.numpy() for a in A], [c.numpy() for c in b], eps=args.eps, rmax=args.rmax)
        dt = time.perf_counter() - t0
        kind = "port"
    return {"value": 1.0 / dt, "unit": "solves/s", "cores": threads, "kind": kind,
            "sample": "1 full solve of the same workload (%.2f s)" % dt}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: independent TT systems shard across GPUs with no "
              "communication, so --gpus %d is not run here." % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torchtt_amd as tt
    from torchtt_amd import workloads as W
    from torchtt_amd import _lib

    lib = _lib.load()
    dev = torch.device("cuda", 0)
    A, b = W.laplace_potential(args.d, args.n, seed=0)
    Ag = tt.TT([c.to(dev) for c in A])
    bg = tt.TT([c.to(dev) for c in b])
    for _ in range(args.warmup):
        torch.manual_seed(0)
        tt.solvers.amen_solve(Ag, bg, eps=args.eps, rmax=args.rmax, precision=args.precision)
    torch.cuda.synchronize()
    _lib.check("profile_enable", lib.ttamen_profile_enable(1))
    lat, sols, stats = [], [], []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        torch.manual_seed(0)
        t0 = time.perf_counter()
        x = tt.solvers.amen_solve(Ag, bg, eps=args.eps, rmax=args.rmax, precision=args.precision)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        sols.append(x)
        stats.append(dict(tt.solvers.last_stats))
    torch.cuda.synchronize()
    total = time.perf_counter() - t_start
    prof = (ctypes.c_double * 4)()
    _lib.check("profile_read", lib.ttamen_profile_read(prof))
    lib.ttamen_profile_enable(0)
    residuals = [tt.solvers.residual(Ag, x, bg) for x in sols]
    n_mv, ms_mv, fl_mv, by_mv = prof[0], prof[1], prof[2], prof[3]
    avg_s = (ms_mv / 1e3) / max(n_mv, 1)
    achieved_tf = (fl_mv / max(n_mv, 1)) / avg_s / 1e12 if n_mv else 0.0
    peak = PEAK_TFLOPS[args.precision]
    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": METRIC,
        "value": args.steps / total,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64-refinement",
        "data": "synthetic",
        "config": {"workload": "AMEn solve, Laplacian + potential, d=%d, n=%d, TT rank 4, random rank-4 rhs, "
                               "rank cap %d, eps %g, %s" % (args.d, args.n, args.rmax, args.eps, args.precision),
                   "d": args.d, "n": args.n, "rmax": args.rmax, "eps": args.eps, "precision": args.precision},
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "residual_max": max(residuals),
        "all_residuals_le_1e-8": bool(max(residuals) <= 1e-8),
        "sweeps": stats[-1]["sweeps"],
        "gmres_iters": stats[-1]["gmres_iters"],
        "max_rank": stats[-1]["max_rank"],
        "roofline": {"kernel": "local_matvec (3 MFMA GEMM launches per application)", "bound": "mfma",
                     "achieved": achieved_tf, "peak": peak, "unit": "TFLOP/s", "frac": achieved_tf / peak,
                     "traffic": None, "applications": int(n_mv), "avg_us": avg_s * 1e6,
                     "flops_per_application": fl_mv / max(n_mv, 1),
                     "compulsory_bytes_per_application": by_mv / max(n_mv, 1)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(A, b, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
