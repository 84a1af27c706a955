"""Benchmark of the AMEn linear solver on one MI355X.

Workloads (--workload):
  batch      (default) 256 independent anisotropic-diffusion systems: d ~ U{4..12},
             n ~ U{8..64}, coefficients log-uniform in [0.1, 10] per mode, operator TT rank 2-4,
             random rank-1..4 N(0,1) right-hand sides with ||b|| = 1, rank caps U{16..64},
             eps 1e-8. A step is one pass over all 256 systems with the batched solver in
             --precision (default fp32: FP32 local GMRES on the matrix cores inside FP64
             refinement): the systems run as one cohort whose kernels are merged into grouped
             launches at every host synchronisation point. The plain FP64 solver, one system at a
             time, is measured on a sample of the same systems for the comparison.
  potential  one system: d=10, n=64, Laplacian + potential (TT rank 4), random rank-4 rhs,
             rank cap 80, eps 1e-8; a step is one solve.
  precision  one system: d=12, n=64, rank cap 128, solved in FP64 and in the two mixed
             precisions (FP32 products; BF16 products with FP32 accumulation); the error of each
             mixed solution vs the FP64 one is reported.
  serve      Poisson arrivals at --rate requests/s of systems drawn as in `batch`, for
             --duration seconds, served by the continuous-batching loop; reports p50/p99
             latency (completion - arrival) and solves/s.

`value` counts only solves whose relative residual ||Ax - b|| / ||b|| is <= 1e-8 (checked
for every result in FP64 TT arithmetic), divided by the wall time of the timed steps, with
inputs resident in HBM. The roofline is that of the local operator, timed with HIP events
inside the library on the solver's streams (FP32 applications inside GMRES for fp32, FP64
otherwise) during a separate single-system solve, so that the timed steps run without the
events; beside it the batch's merged GEMM launches (MFMA) and the GMRES orthogonalisation (HBM).
The CPU baseline is the reference's own C++ solver on this host's cores.
"""
import os

# Hardware queues per process: the batched solver runs 4 cohort lanes on their own streams beside
# torch's; with the runtime's default of 4 queues two of them would share one and serialize (measured:
# 4 lanes on 4 queues 20.1 solves/s, on 8 queues 31.1; 3 lanes 30.4 either way).
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

import argparse  # noqa: E402
import ctypes  # noqa: E402
import glob  # noqa: E402
import importlib.util  # noqa: E402
import json  # noqa: E402
import sys  # noqa: E402
import time  # noqa: E402

import numpy as np  # noqa: E402
import torch  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/sec at rel. residual ≤1e-8, mixed vs FP64 plain; p50/p99 solve latency"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
PEAK_HBM_GBS = 8000.0
RES_BAR = 1e-8


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default="batch", choices=["batch", "potential", "precision", "serve", "solo"])
    p.add_argument("--precision", default=None, choices=["fp64", "fp32", "fp32-phi", "bf16"])
    p.add_argument("--systems", type=int, default=256)
    p.add_argument("--workers", type=int, default=8, help="host threads running the systems' host code")
    p.add_argument("--slots", type=int, default=256, help="serve: requests running at once")
    p.add_argument("--lanes", type=int, default=4, help="concurrent cohorts (streams) of the batched solver")
    p.add_argument("--window", type=int, default=128, help="cohort: grouped launches per round")
    p.add_argument("--max-age", type=int, default=64, help="cohort: merge steps a system may wait")
    p.add_argument("--gmres-poll", type=int, default=8, help="cohort: GMRES iterations between convergence polls")
    p.add_argument("--idle-pct", type=int, default=15, help="cohort: idle weight share that ends a round")
    p.add_argument("--plain-sample", type=int, default=32)
    p.add_argument("--eps", type=float, default=1e-8)
    p.add_argument("--rate", type=float, default=200.0)
    p.add_argument("--duration", type=float, default=2.0)
    p.add_argument("--serve-duration", type=float, default=10.0,
                   help="batch: seconds of Poisson arrivals in the serving leg after the batch (0: no leg)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--matvec-mode", type=int, default=0, choices=[0, 1, 2],
                   help="local operator: 0 GEMM + fused stages 2-3, 1 one fused kernel, 2 three GEMMs")
    p.add_argument("--autotune", type=int, default=0, choices=[0, 1],
                   help="GEMM tile autotuner for single-system solves (1: tune each shape bucket once)")
    a = p.parse_args()
    if a.precision is None:
        a.precision = "fp64" if a.workload == "potential" else "fp32"
    return a


def load_ref():
    so = glob.glob(os.path.join(ROOT, "oracle", "_ref", "torchttcpp*.so"))
    if not so:
        return None
    spec = importlib.util.spec_from_file_location("torchttcpp", so[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def cpu_baseline(items, eps, args, budget_s=25.0, rng_states=None, gpu=None):
    """The reference's own C++ AMEn (oracle/_ref, built from its sources) on this host's cores,
    solving the given (A, b, rmax) CPU systems one after another until `budget_s` is spent;
    the CPU oracle restatement if that build is absent.

    rng_states[i] (optional): torch's CPU generator state from which the GPU run drew system i's
    residual-proxy cores; restored before the reference's solve, so that the reference draws the
    same cores (cpp/amen_solve.h:174-176) and solves the same problem the GPU solved.
    gpu (optional): per-system (residual, ranks) of the GPU's results, recorded beside the
    reference's own residual and ranks (computed outside the timed solves)."""
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    torch.set_num_threads(threads)
    ref = load_ref()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import amen_oracle as O
    done, t_solve, sols = 0, 0.0, []
    for i, (A, b, rmax) in enumerate(items):
        d = len(A)
        rA = [1] + [a.shape[3] for a in A]
        rb = [1] + [c.shape[2] for c in b]
        if rng_states is not None:
            torch.set_rng_state(rng_states[i])
        else:
            torch.manual_seed(0)
        t0 = time.perf_counter()
        if ref is not None:
            x = ref.amen_solve(A, b, [], [a.shape[2] for a in A], rA, rb, [1] * (d + 1), 22, eps, rmax, 500,
                               4, 0, 40, 2, False, 0)
            x = [c.numpy() for c in x]
        else:
            x = O.amen_solve([a.numpy() for a in A], [c.numpy() for c in b], eps=eps, rmax=rmax)
        t_solve += time.perf_counter() - t0
        sols.append(x)
        done += 1
        if t_solve > budget_s:
            break
    out = {"value": done / t_solve, "unit": "solves/s", "cores": threads,
           "kind": "reference" if ref is not None else "port",
           "sample": "first %d system(s) of the workload solved one after another (%.1f s)" % (done, t_solve)}
    # accuracy of each reference solve (FP64 TT residual by the oracle's QR-based norm), beside the GPU's
    res = [O.tt_residual([a.numpy() for a in A], x, [c.numpy() for c in b]) for (A, b, _), x in zip(items, sols)]
    out["residuals"] = res
    out["solved_le_1e-8"] = int(sum(r <= RES_BAR for r in res))
    out["max_ranks"] = [max(c.shape[2] for c in x) for x in sols]
    if gpu is not None:
        g_res, g_ranks = gpu
        out["gpu_residuals"] = [float(v) for v in g_res[:done]]
        out["gpu_solved_le_1e-8"] = int(sum(v <= RES_BAR for v in g_res[:done]))
        out["final_ranks_equal_to_gpu"] = int(sum(list(g_ranks[i]) == [1] + [c.shape[2] for c in x]
                                                  for i, x in enumerate(sols)))
        out["stalled_in_both"] = int(sum(g_res[i] > RES_BAR and res[i] > RES_BAR for i in range(done)))
        out["stalled_on_gpu_only"] = int(sum(g_res[i] > RES_BAR and res[i] <= RES_BAR for i in range(done)))
    return out


def pmc_traffic(precision):
    """HBM bytes per local-matvec application from the committed PMC summary (FETCH_SIZE +
    WRITE_SIZE, scaled by a calibration copy), for the FP32 matvec it was measured on."""
    path = os.path.join(ROOT, "profiles", "matvec_fp32_largest_batch_system_pmc.md")
    if precision != "fp32" or not os.path.exists(path):
        return None, None
    txt = open(path).read()
    i = txt.find("<!-- pmc-json")
    if i < 0:
        return None, None
    blob = txt[i + len("<!-- pmc-json"):txt.index("-->", i)]
    return json.loads(blob)["traffic_bytes_per_application"], os.path.relpath(path, ROOT)


MATVEC_MODE = [0]  # the local-operator form of this run (--matvec-mode)
MATVEC_FORM = {0: "a stage-1 MFMA GEMM + one fused stages-2/3 MFMA kernel per application",
               1: "one fused kernel per application", 2: "3 MFMA GEMM launches per application"}


def roofline(prof, precision):
    mode = MATVEC_MODE[0]
    n64, ms64, fl64, by64, n32, ms32, fl32, by32 = list(prof)
    if precision == "bf16" and n32 > 0:
        n, ms, fl, by, peak = n32, ms32, fl32, by32, PEAK_TFLOPS["bf16"]
        kern = ("local matvec BF16 products, FP32 accumulation (3 MFMA GEMM launches per application), "
                "Arnoldi steps that ran")
    elif precision in ("fp32", "fp32-phi") and n32 > 0:
        n, ms, fl, by, peak = n32, ms32, fl32, by32, PEAK_TFLOPS["fp32"]
        kern = "local matvec FP32 (%s), Arnoldi steps that ran" % MATVEC_FORM[mode]
    else:
        n, ms, fl, by, peak = n64, ms64, fl64, by64, PEAK_TFLOPS["fp64"]
        kern = "local matvec FP64 (%s)" % MATVEC_FORM[mode]
    avg_s = (ms / 1e3) / max(n, 1)
    ach = (fl / max(n, 1)) / avg_s / 1e12 if n else 0.0
    traffic, src = pmc_traffic(precision)
    return {"kernel": kern, "bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s",
            "frac": ach / peak, "traffic": traffic, "traffic_source": src,
            "applications": int(n), "avg_us": avg_s * 1e6,
            "flops_per_application": fl / max(n, 1),
            "flops_note": ("the operator-core contraction counted over the core's nonzero 16x16 tiles (the steps "
                           "the kernel runs), the two interface contractions in full"
                           if mode == 0 and precision != "bf16" else "the dense contraction (this form multiplies "
                           "every operator tile)"),
            "compulsory_bytes_per_application": by / max(n, 1)}


def base_record(args, value, total, config, dtype):
    return {"metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": total / max(args.steps, 1) * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


def dtype_of(p):
    return {"fp64": "f64", "fp32": "f32-inner+f64-refinement", "fp32-phi": "f32-inner+f32-interfaces+f64-refinement",
            "bf16": "bf16-products+f32-acc+f64-refinement"}[p]


def system_cost(sy):
    """Work estimate of a batch system: sum over cores of the local problem size at its rank cap."""
    n, d, r = sy["n"], sy["d"], sy["rmax"] + 4
    return d * n * r * r


def ortho_roofline(lib, _lib, precision):
    """HBM roofline of the GMRES orthogonalisation per Arnoldi step, from the library's HIP events and
    algorithmic bytes, read after a profiled single-system solve: FP32 basis, CGS dots + update + close
    ((2k + 7) N w bytes at step k); FP64 basis, update + close ((k + 5) N w: the first pass's dots run inside
    the local operator's stages-2/3 kernel wherever that kernel's grid allows, and are not timed here)."""
    o = (ctypes.c_double * 3)()
    _lib.check("profile_read_ortho", lib.ttamen_profile_read_ortho(o))
    steps, ms, by = list(o)
    if steps <= 0 or ms <= 0:
        return None
    ach = by / (ms / 1e3) / 1e9
    if precision == "fp64":
        kern = ("GMRES orthogonalisation: update (+ second-pass dots) and close launches of an Arnoldi step, FP64 "
                "basis (the first pass's dots fused into the local operator's kernel)")
    else:
        kern = ("GMRES orthogonalisation: CGS dots, update (+ second-pass dots) and close launches of an Arnoldi "
                "step (FP32 basis)")
    return {"kernel": kern,
            "bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": ach / PEAK_HBM_GBS,
            "steps": int(steps), "avg_us": ms * 1e3 / steps, "bytes_per_step": by / steps}


def solo_roofline(tt, lib, _lib, item, z, precision, eps):
    """Local-matvec roofline from HIP events around every application in one single-system
    solve (on the solver's stream)."""
    A, b, rm = item
    _lib.check("profile_enable", lib.ttamen_profile_enable(1))
    tt.solvers.amen_solve(A, b, rmax=rm, eps=eps, _z0=z, precision=precision)
    torch.cuda.synchronize()
    prof = (ctypes.c_double * 8)()
    _lib.check("profile_read", lib.ttamen_profile_read(prof))
    orth = ortho_roofline(lib, _lib, precision)
    lib.ttamen_profile_enable(0)
    out = roofline(prof, precision)
    return out, orth


def run_batch(args, tt, W, lib, _lib):
    dev = torch.device("cuda", 0)
    systems = W.batch_systems(args.systems, seed=0)
    items = [(tt.TT([c.to(dev) for c in s["A"]]), tt.TT([c.to(dev) for c in s["b"]]), s["rmax"]) for s in systems]
    torch.manual_seed(0)
    z0, rng_states = [], []
    for it in items:  # the generator state before each system's draws (the CPU baseline replays them)
        rng_states.append(torch.get_rng_state())
        z0.append(tt.solvers._draw_z(it[1].N, 4, 0, "cpu"))
    kw = dict(eps=args.eps, workers=args.workers)
    for _ in range(args.warmup):
        tt.solvers.amen_solve_batch(items[:16], precision=args.precision, z0=z0[:16], **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sols = None
    for _ in range(args.steps):
        sols = tt.solvers.amen_solve_batch(items, precision=args.precision, z0=z0, **kw)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    stats = tt.solvers.last_batch_stats
    cohort = {k: round(v, 4) for k, v in tt.solvers.last_cohort_stats.items()}
    res = np.array([tt.solvers.residual(it[0], x, it[1]) for it, x in zip(items, sols)])
    ok = int(np.sum(res <= RES_BAR))
    lat_ms = np.array([s["elapsed_s"] for s in stats]) * 1e3
    value = ok * args.steps / total

    # the same batch with FP64 local solves (one pass)
    other = "fp64" if args.precision == "fp32" else "fp32"
    t1 = time.perf_counter()
    sols_o = tt.solvers.amen_solve_batch(items, precision=other, z0=z0, **kw)
    torch.cuda.synchronize()
    t_o = time.perf_counter() - t1
    res_o = np.array([tt.solvers.residual(it[0], x, it[1]) for it, x in zip(items, sols_o)])
    ok_o = int(np.sum(res_o <= RES_BAR))

    # plain FP64 solver, one system at a time, on a sample of the same systems
    ns = min(args.plain_sample, len(items))
    t2 = time.perf_counter()
    plain = [tt.solvers.amen_solve(A, b, rmax=rm, eps=args.eps, _z0=z) for (A, b, rm), z in zip(items[:ns], z0[:ns])]
    torch.cuda.synchronize()
    tplain = time.perf_counter() - t2
    res_plain = np.array([tt.solvers.residual(it[0], x, it[1]) for it, x in zip(items[:ns], plain)])
    ok_plain = int(np.sum(res_plain <= RES_BAR))
    plain_rate = ok_plain / tplain if ns else None
    err = [float((x - y).norm() / y.norm()) for x, y in zip(sols[:ns], plain)]
    costs = [system_cost(sy) for sy in systems]
    sample_share = float(sum(costs[:ns]) / sum(costs))

    # the merged GEMM launches of the batch: a repeat of the timed pass with HIP events around each
    gp = (ctypes.c_double * 6)()
    _lib.check("cohort_gemm_profile", lib.ttamen_cohort_gemm_profile(1, None))
    tt.solvers.amen_solve_batch(items, precision=args.precision, z0=z0, **kw)
    torch.cuda.synchronize()
    _lib.check("cohort_gemm_profile", lib.ttamen_cohort_gemm_profile(0, gp))
    g_launch, g_ms, g64, g32, gbf, g_items = list(gp)
    roof_b = None
    if g_launch > 0 and g_ms > 0:
        t_g = g_ms / 1e3
        busy = (g64 / (PEAK_TFLOPS["fp64"] * 1e12) + g32 / (PEAK_TFLOPS["fp32"] * 1e12)
                + gbf / (PEAK_TFLOPS["bf16"] * 1e12)) / t_g
        ach = (g64 + g32 + gbf) / t_g / 1e12
        roof_b = {"kernel": "k_mega_gemm: the batch's merged GEMM launches (items of many systems: FP32 local-matvec "
                            "stages, FP64 refinement residuals, interface updates, QR Gram products)",
                  "bound": "mfma", "achieved": ach, "peak": ach / busy if busy > 0 else None, "unit": "TFLOP/s",
                  "frac": busy, "launches": int(g_launch), "avg_us": g_ms * 1e3 / g_launch,
                  "items_per_launch": g_items / g_launch, "flops_f64": g64, "flops_f32": g32, "flops_bf16": gbf,
                  "peak_note": "flop-weighted mix of the FP64 (78.6) and FP32 (157.3) MFMA peaks; frac = time the "
                               "items' flops need at peak / measured launch time",
                  "measured_on": "a repeat of the timed pass with HIP events around every merged GEMM launch"}

    big = int(np.argmax([system_cost(sy) for sy in systems]))
    roof, roof_orth = solo_roofline(tt, lib, _lib, items[big], z0[big], args.precision, args.eps)
    roof["measured_on"] = ("single solve of the largest batch system (index %d: d=%d, n=%d, rank cap %d); "
                           "inside the batch the GEMMs are items of merged launches"
                           % (big, systems[big]["d"], systems[big]["n"], systems[big]["rmax"]))

    cfg = {"workload": "batched AMEn: %d anisotropic diffusion systems, d~U{4..12}, n~U{8..64}, op rank 2-4, "
                       "rhs rank 1-4, rank caps U{16..64}, eps %g, %s local solves, cohort execution"
                       % (len(items), args.eps, args.precision),
           "systems": len(items), "precision": args.precision, "host_threads": args.workers, "eps": args.eps}
    out = base_record(args, value, total, cfg, dtype_of(args.precision))
    out.update({
        "systems_per_s_all": len(items) * args.steps / total,
        "solved_le_1e-8": ok, "systems": len(items),
        "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
        "latency_note": "per-system solve time inside the batch (all admitted at t=0)",
        "residual_max_of_solved": float(res[res <= RES_BAR].max()) if ok else None,
        "batched_" + other: {"time_s": t_o, "solves_per_s": ok_o / t_o, "solved_le_1e-8": ok_o},
        "plain_fp64_sequential": {"sample_systems": ns, "time_s": tplain, "solves_per_s": plain_rate,
                                  "solved_le_1e-8": ok_plain,
                                  "sample_cost_share": sample_share,
                                  "cost_share_note": "the sample's share of the batch's work estimate (sum over "
                                                     "systems of d n (rmax + 4)^2); %d of %d systems"
                                                     % (ns, len(items))},
        "speedup_vs_plain_fp64_sequential": value / plain_rate if plain_rate else None,
        "error_vs_plain_fp64_max": max(err) if err else None,
        "error_vs_plain_fp64_median": float(np.median(err)) if err else None,
        "roofline": roof,
        "roofline_batched_gemm": roof_b,
        "roofline_orthogonalisation": roof_orth,
        "cohort": cohort,
    })
    if args.serve_duration > 0:
        out["serve"] = serve_leg(args, tt, W, args.serve_duration)
    if not args.no_cpu_baseline:
        cpu_items = [(s["A"], s["b"], s["rmax"]) for s in systems]
        g_ranks = [x.R for x in plain]
        out["cpu_baseline"] = cpu_baseline(cpu_items[:ns], args.eps, args, rng_states=rng_states[:ns],
                                           gpu=(res_plain, g_ranks))
        out["cpu_baseline"]["gpu_result_compared"] = "plain FP64 GPU solve of the same system and draws"
    return out


def serve_leg(args, tt, W, duration):
    """Latency under arrivals: Poisson arrivals at --rate requests/s for `duration` seconds of systems
    drawn as the batch's, submitted at their arrival times to the online server (serving.Server) in
    --precision while earlier requests run."""
    dev = torch.device("cuda", 0)
    arrivals = tt.serving.poisson_arrivals(args.rate, duration, seed=1)
    systems = W.batch_systems(len(arrivals), seed=7)
    items = [(tt.TT([c.to(dev) for c in s["A"]]), tt.TT([c.to(dev) for c in s["b"]]), s["rmax"]) for s in systems]
    torch.manual_seed(0)
    z0 = [tt.solvers._draw_z(it[1].N, 4, 0, "cpu") for it in items]
    lib = tt._lib.load()
    need = 0
    for A, b, rm in items:
        need = max(need, lib.ttamen_amen_solve_workspace_size(len(b.N), tt._lib.i64_array(b.N), tt._lib.i64_array(A.R),
                                                              tt._lib.i64_array(b.R), None, rm, 4, 0, 40, 500,
                                                              {"fp64": 0, "fp32": 1, "bf16": 2, "fp32-phi": 3}[args.precision]))
    print("serve leg: %d requests over %g s" % (len(items), duration), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    with tt.serving.Server(slots=args.slots, memory_bytes=args.slots * (need + 256), eps=args.eps,
                           precision=args.precision) as srv:
        t0 = time.perf_counter()
        reqs = []
        for (A, b, rm), z, a in zip(items, z0, arrivals):
            wait = t0 + a - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            reqs.append(srv.submit(A, b, rmax=rm, z0=z))
        sols = [r.result() for r in reqs]
        times = np.array([r.times() for r in reqs])
        total = time.perf_counter() - t0
    res = np.array([tt.solvers.residual(it[0], x, it[1]) for it, x in zip(items, sols)])
    lat = (times[:, 2] - times[:, 0]) * 1e3
    span = float(times[:, 2].max() - times[:, 0].min())
    late = np.array([t - (t0 + a) for t, a in zip(times[:, 0] - times[0, 0] + t0, arrivals)])
    return {"api": "online server (ttamen_server_submit at each arrival time, ttamen_server_wait)",
            "rate_req_s": args.rate, "duration_s": duration, "requests": len(items), "slots": args.slots,
            "precision": args.precision, "makespan_s": total, "drain_s": span,
            "served_per_s": len(items) / span, "solves_per_s_le_1e-8": int(np.sum(res <= RES_BAR)) / span,
            "solved_le_1e-8": int(np.sum(res <= RES_BAR)),
            "serve_p50_ms": float(np.percentile(lat, 50)), "serve_p99_ms": float(np.percentile(lat, 99)),
            "queue_wait_p50_ms": float(np.percentile((times[:, 1] - times[:, 0]) * 1e3, 50)),
            "submission_lag_p99_ms": float(np.percentile(np.abs(late) * 1e3, 99)),
            "note": "latency = completion - submission, with submissions at the Poisson arrival times; "
                    "drain_s from the first submission to the last completion"}


def run_potential(args, tt, W, lib, _lib):
    dev = torch.device("cuda", 0)
    A, b = W.laplace_potential(10, 64, seed=0)
    Ag = tt.TT([c.to(dev) for c in A])
    bg = tt.TT([c.to(dev) for c in b])
    for _ in range(args.warmup):
        torch.manual_seed(0)
        tt.solvers.amen_solve(Ag, bg, eps=args.eps, rmax=80, precision=args.precision)
    torch.cuda.synchronize()
    lat, sols, stats = [], [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        torch.manual_seed(0)
        t0 = time.perf_counter()
        x = tt.solvers.amen_solve(Ag, bg, eps=args.eps, rmax=80, precision=args.precision)
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        sols.append(x)
        stats.append(dict(tt.solvers.last_stats))
    torch.cuda.synchronize()
    total = time.perf_counter() - t_start
    # rooflines from one more solve with the library's timing events on (kept out of the timed steps)
    _lib.check("profile_enable", lib.ttamen_profile_enable(1))
    torch.manual_seed(0)
    tt.solvers.amen_solve(Ag, bg, eps=args.eps, rmax=80, precision=args.precision)
    torch.cuda.synchronize()
    prof = (ctypes.c_double * 8)()
    _lib.check("profile_read", lib.ttamen_profile_read(prof))
    roof_orth = ortho_roofline(lib, _lib, args.precision)
    lib.ttamen_profile_enable(0)
    res = np.array([tt.solvers.residual(Ag, x, bg) for x in sols])
    ok = int(np.sum(res <= RES_BAR))
    lat_ms = np.array(lat) * 1e3
    cfg = {"workload": "AMEn solve, Laplacian + potential, d=10, n=64, TT rank 4, random rank-4 rhs, "
                       "rank cap 80, eps %g, %s" % (args.eps, args.precision),
           "d": 10, "n": 64, "rmax": 80, "eps": args.eps, "precision": args.precision}
    out = base_record(args, ok / total, total, cfg, dtype_of(args.precision))
    out.update({"p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
                "residual_max": float(res.max()), "solved_le_1e-8": ok, "sweeps": stats[-1]["sweeps"],
                "gmres_iters": stats[-1]["gmres_iters"], "max_rank": stats[-1]["max_rank"],
                "roofline": roofline(prof, args.precision), "roofline_orthogonalisation": roof_orth})
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline([(A, b, 80)], args.eps, args, budget_s=0.0)
    return out


def run_precision(args, tt, W, lib, _lib):
    """d=12, n=64, Laplacian + potential (TT rank 4), random rank-4 rhs, rank cap 128, eps 1e-8:
    one FP64 solve and two mixed-precision solves (FP32 local GMRES on the local operator with
    FP32 products, and with BF16 products and FP32 accumulation, inside FP64 refinement) from the
    same residual-proxy draws; error of each mixed solution vs the FP64 one, residual of each."""
    dev = torch.device("cuda", 0)
    A, b = W.laplace_potential(12, 64, seed=0)
    Ag = tt.TT([c.to(dev) for c in A])
    bg = tt.TT([c.to(dev) for c in b])
    torch.manual_seed(0)
    z = tt.solvers._draw_z(bg.N, 4, 0, "cpu")
    rows = {}
    for prec in ("fp64", "fp32", "fp32-phi", "bf16"):
        for _ in range(args.warmup):
            tt.solvers.amen_solve(Ag, bg, eps=args.eps, rmax=128, precision=prec, _z0=z)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = tt.solvers.amen_solve(Ag, bg, eps=args.eps, rmax=128, precision=prec, _z0=z)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        # roofline from one more solve with the library's timing events on
        _lib.check("profile_enable", lib.ttamen_profile_enable(1))
        tt.solvers.amen_solve(Ag, bg, eps=args.eps, rmax=128, precision=prec, _z0=z)
        torch.cuda.synchronize()
        prof = (ctypes.c_double * 8)()
        _lib.check("profile_read", lib.ttamen_profile_read(prof))
        lib.ttamen_profile_enable(0)
        st = dict(tt.solvers.last_stats)
        rows[prec] = {"x": x, "time_s": dt, "residual": tt.solvers.residual(Ag, x, bg), "sweeps": st["sweeps"],
                      "gmres_iters": st["gmres_iters"], "max_rank": st["max_rank"],
                      "refinements": st.get("refinements", 0), "roofline": roofline(prof, prec)}
    x64, x32, xbf = rows["fp64"].pop("x"), rows["fp32"].pop("x"), rows["bf16"].pop("x")
    xphi = rows["fp32-phi"].pop("x")
    err = float((x32 - x64).norm() / x64.norm())
    err_bf = float((xbf - x64).norm() / x64.norm())
    err_phi = float((xphi - x64).norm() / x64.norm())
    ok = sum(1 for r in rows.values() if r["residual"] <= RES_BAR)
    total = sum(r["time_s"] for r in rows.values())
    cfg = {"workload": "AMEn solve, Laplacian + potential, d=12, n=64, TT rank 4, random rank-4 rhs, rank cap "
                       "128, eps %g: FP64, FP32-inner and BF16-products-inner, both + FP64-refinement" % args.eps,
           "d": 12, "n": 64, "rmax": 128, "eps": args.eps}
    out = base_record(args, 1.0 / rows["fp32"]["time_s"] if rows["fp32"]["residual"] <= RES_BAR else 0.0,
                      total, cfg, dtype_of("fp32"))
    out.update({"fp64": rows["fp64"], "fp32": rows["fp32"], "fp32-phi": rows["fp32-phi"], "bf16": rows["bf16"],
                "error_fp32_vs_fp64": err, "error_fp32_phi_vs_fp64": err_phi,
                "error_bf16_vs_fp64": err_bf, "solved_le_1e-8": ok,
                "roofline": rows["fp32"]["roofline"]})
    return out


def run_solo(args, tt, W, lib, _lib):
    """The single solve the batch roofline is measured on (the largest batch system), preceded
    by a 64 MiB device copy that calibrates the HBM counters of a PMC run of this command."""
    dev = torch.device("cuda", 0)
    src = torch.ones(8 << 20, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    systems = W.batch_systems(args.systems, seed=0)
    big = int(np.argmax([system_cost(sy) for sy in systems]))
    sy = systems[big]
    item = (tt.TT([c.to(dev) for c in sy["A"]]), tt.TT([c.to(dev) for c in sy["b"]]), sy["rmax"])
    torch.manual_seed(0)
    # the same residual-proxy draws as in the batched run (drawn system by system)
    z0 = [tt.solvers._draw_z([c.shape[1] for c in s["b"]], 4, 0, "cpu") for s in systems[:big + 1]][big]
    t0 = time.perf_counter()
    roof, roof_orth = solo_roofline(tt, lib, _lib, item, z0, args.precision, args.eps)
    total = time.perf_counter() - t0
    cfg = {"workload": "single solve of the largest batch system (index %d: d=%d, n=%d, rank cap %d), %s"
                       % (big, sy["d"], sy["n"], sy["rmax"], args.precision), "calibration_copy_bytes": 64 << 20}
    out = base_record(args, 1.0 / total, total, cfg, dtype_of(args.precision))
    out["roofline"] = roof
    out["roofline_orthogonalisation"] = roof_orth
    return out


def run_serve(args, tt, W, lib, _lib):
    dev = torch.device("cuda", 0)
    arrivals = tt.serving.poisson_arrivals(args.rate, args.duration, seed=1)
    print("serve: building %d requests" % len(arrivals), file=sys.stderr, flush=True)
    systems = W.batch_systems(len(arrivals), seed=7)
    items = [(tt.TT([c.to(dev) for c in s["A"]]), tt.TT([c.to(dev) for c in s["b"]]), s["rmax"]) for s in systems]
    torch.manual_seed(0)
    z0 = [tt.solvers._draw_z(it[1].N, 4, 0, "cpu") for it in items]
    if args.warmup:
        tt.solvers.amen_solve_batch(items[:8], eps=args.eps, precision=args.precision, z0=z0[:8], workers=4)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sols, info = tt.serving.serve(items, arrivals, slots=args.slots, eps=args.eps, precision=args.precision,
                                  z0=z0)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    print("serve: %d requests served in %.1f s; checking residuals" % (len(items), total), file=sys.stderr,
          flush=True)
    res = []
    for i, (it, x) in enumerate(zip(items, sols)):
        res.append(tt.solvers.residual(it[0], x, it[1]))
        if (i + 1) % 2000 == 0:
            print("serve: %d residuals checked" % (i + 1), file=sys.stderr, flush=True)
    res = np.array(res)
    ok = int(np.sum(res <= RES_BAR))
    lat = info["latency_s"] * 1e3
    cfg = {"workload": "served AMEn: Poisson arrivals at %g req/s for %g s (%d requests) of systems drawn as in "
                       "the batched workload, %d slots, %s, eps %g"
                       % (args.rate, args.duration, len(items), args.slots, args.precision, args.eps),
           "rate": args.rate, "duration_s": args.duration, "requests": len(items), "slots": args.slots}
    out = base_record(args, ok / total, total, cfg, dtype_of(args.precision))
    out.update({"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "solved_le_1e-8": ok, "requests": len(items), "makespan_s": total,
                "queue_wait_p50_ms": float(np.percentile((info["start_s"] - info["arrival_s"]) * 1e3, 50))})
    return out


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
    _lib.check("cohort_set_params", lib.ttamen_cohort_set_params(args.window, args.max_age, args.gmres_poll,
                                                                 args.lanes, args.idle_pct))
    lib.ttamen_gemm_autotune(args.autotune)
    lib.ttamen_matvec_mode(args.matvec_mode)
    MATVEC_MODE[0] = args.matvec_mode
    fn = {"batch": run_batch, "potential": run_potential, "precision": run_precision,
          "serve": run_serve, "solo": run_solo}[args.workload]
    out = fn(args, tt, W, lib, _lib)
    if args.autotune:
        st = (ctypes.c_double * 4)()
        _lib.check("gemm_autotune_stats", lib.ttamen_gemm_autotune_stats(st))
        out["gemm_autotune"] = {"buckets_tuned": st[0], "launches_from_cache": st[1], "tuning_s": st[2],
                                "launches_off_rule": st[3]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
