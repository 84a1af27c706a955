"""AMEn linear solver for TT systems A x = b -- drop-in for `torchtt.solvers.amen_solve`.

Same signature, argument checks, errors and TT return as the reference
(torchtt/solvers.py:211-287). The solve itself runs in libttamen's HIP kernels on the
GPU (include/ttamen.h: ttamen_amen_solve, which replaces torchttcpp.amen_solve,
cpp/amen_solve.h:122-549). There is no CPU fallback: without the library or a ROCm GPU
the call raises.
"""
import ctypes
import re
import sys

import numpy as np
import torch

from ._tt import TT
from .errors import InvalidArguments, IncompatibleTypes, ShapeMismatch
from . import _lib

#: statistics of the most recent solve (dict), see include/ttamen.h: ttamen_stats
last_stats = None

_PRECISIONS = {"fp64": 0, "fp32": 1, "bf16": 2, "fp32-phi": 3}


def cpp_enabled():
    """True when the native solver library is built (reference solvers.py:26-33)."""
    try:
        _lib.load()
        return True
    except ImportError:
        return False


def _solver_device(A):
    dev = A.cores[0].device
    if dev.type == "cuda":
        return dev
    if not torch.cuda.is_available():
        raise RuntimeError("torchtt_amd.solvers.amen_solve runs on a ROCm GPU; none is available.")
    return torch.device("cuda", torch.cuda.current_device())


def _require_real(*tts):
    """The native sweeps compute in real FP64: refuse complex cores (a cast would drop their
    imaginary parts)."""
    for t in tts:
        if t is not None and any(c.is_complex() for c in t.cores):
            raise InvalidArguments('Complex cores are not supported: the solver computes in real FP64.')


def _check_system(A, b):
    if not (isinstance(A, TT) and isinstance(b, TT)):
        raise InvalidArguments('A and b must be TT instances.')
    if not (A.is_ttm and not b.is_ttm):
        raise IncompatibleTypes('A must be TT-matrix and b must be vector.')
    if A.M != A.N:
        raise ShapeMismatch('A is not quadratic.')
    if A.N != b.N:
        raise ShapeMismatch('Dimension mismatch.')
    _require_real(A, b)


def _draw_z(N, kickrank, kick2, device):
    """Residual-proxy initial cores, drawn as the reference draws them (cpp/amen_solve.h:169-176)."""
    d = len(N)
    rzc = kickrank + kick2
    if rzc <= 0:
        return None
    rz = [1] + [rzc] * (d - 1) + [1]
    return [torch.randn((rz[i], N[i], rz[i + 1]), dtype=torch.float64, device=device) for i in range(d)]


def amen_solve(A, b, nswp=22, x0=None, eps=1e-10, rmax=32768, max_full=500, kickrank=4, kick2=0,
               trunc_norm='res', local_solver=1, local_iterations=40, resets=2, verbose=False,
               preconditioner=None, use_cpp=True, band_diagonal=-1, use_single_precision=False,
               precision=None, seed=None, _z0=None):
    """Solve A x = b in the TT format with AMEn (Dolgov & Savostyanov, SIAM J. Sci. Comput. 36 (2014)).

    Args follow the reference. `trunc_norm`, `local_solver`, `use_cpp` and `band_diagonal`
    are accepted for compatibility; the native path always uses GMRES and the residual-based
    truncation, as the reference's C++ path does. New arguments:
        precision: 'fp64' (default), 'fp32' (FP32 matrix-core local operator with FP64
            iterative refinement), 'fp32-phi' (the same, and the interface updates contracted in
            FP32 on the matrix cores, the interfaces kept in FP64) or 'bf16' (BF16 products and
            FP32 accumulation in the local operator); `use_single_precision=True` selects 'fp32'.
        seed: seed for device-drawn random enrichment columns (only used when kick2 > 0).

    Returns:
        TT: the solution on the device and dtype of A.cores[0].
    """
    global last_stats
    _check_system(A, b)
    if preconditioner is None:
        prec = 0
    elif preconditioner == 'c':
        prec = 1
    elif preconditioner == 'r':
        prec = 2
    else:
        raise InvalidArguments("Invalid preconditioner.")
    if precision is None:
        precision = 'fp32' if use_single_precision else 'fp64'
    if precision not in _PRECISIONS:
        raise InvalidArguments("precision must be one of %s." % sorted(_PRECISIONS))
    if prec != 0 and precision != 'fp64':
        raise InvalidArguments("Local preconditioners are available with FP64 local solves.")

    _require_real(x0)
    lib = _lib.load()
    out_device, out_dtype = A.cores[0].device, A.cores[0].dtype
    dev = _solver_device(A)
    N = list(b.N)
    d = len(N)
    rA, rb = A.R, b.R

    # Residual-proxy initial cores: drawn like the reference draws them (torch.randn with
    # the options of A, cpp/amen_solve.h:174-176), so a seeded run matches it.
    z_cores = _z0 if _z0 is not None else _draw_z(N, kickrank, kick2, out_device)
    if z_cores is not None:
        z_cores = [z.to(device=dev, dtype=torch.float64).contiguous() for z in z_cores]

    with torch.cuda.device(dev):
        Ac = [c.to(device=dev, dtype=torch.float64).contiguous() for c in A.cores]
        bc = [c.to(device=dev, dtype=torch.float64).contiguous() for c in b.cores]
        if x0 is not None:
            xc = [c.to(device=dev, dtype=torch.float64).contiguous() for c in x0.cores]
            rx0 = x0.R
        else:
            xc, rx0 = None, None
        Na, rAa, rba = _lib.i64_array(N), _lib.i64_array(rA), _lib.i64_array(rb)
        rx0a = _lib.i64_array(rx0) if rx0 is not None else None
        cap = None
        while True:
            prev_cap = lib.ttamen_rank_capacity(cap) if cap else None
            try:
                x_cores, ws_bytes = _amen_solve_native(
                    lib, d, N, Na, rAa, rba, Ac, bc, xc, rx0a, z_cores, seed, nswp, eps, rmax, max_full,
                    kickrank, kick2, local_iterations, resets, verbose, prec, precision, dev, out_device,
                    out_dtype)
                break
            except _lib.TTAmenError as e:
                cap = _grown_capacity(e, cap or lib.ttamen_rank_capacity(0))
                if cap is None:
                    raise
            finally:
                if prev_cap is not None:
                    lib.ttamen_rank_capacity(prev_cap)
        last_stats["workspace_bytes"] = int(ws_bytes)
        last_stats["rank_capacity"] = int(cap) if cap else int(lib.ttamen_rank_capacity(0))
    return TT(x_cores)


_RANK_CAPACITY_MAX = 1024
_CAPACITY_MSG = re.compile(r"needs rank (\d+) at interface \d+; ranks above (\d+) are not supported at this "
                           r"rank capacity")


def _grown_capacity(err, current):
    """The rank capacity to solve again with after a sweep outgrew `current` (status 4 at an interface
    planned at the capacity), or None when more capacity would not help."""
    if err.code != 4:
        return None
    m = _CAPACITY_MSG.search(str(err))
    if m is None:
        return None
    need, held = int(m.group(1)), int(m.group(2))
    if held < current:  # that interface was bounded by rmax or by the mode sizes, not the capacity
        return None
    nxt = min(_RANK_CAPACITY_MAX, max(2 * current, need + 16))
    return nxt if nxt > current else None


def _amen_solve_native(lib, d, N, Na, rAa, rba, Ac, bc, xc, rx0a, z_cores, seed, nswp, eps, rmax, max_full,
                       kickrank, kick2, local_iterations, resets, verbose, prec, precision, dev, out_device,
                       out_dtype):
    """One call of ttamen_amen_solve at the calling thread's rank capacity; the solution's cores."""
    global last_stats
    ws_bytes = lib.ttamen_amen_solve_workspace_size(d, Na, rAa, rba, rx0a, int(rmax), int(kickrank),
                                                    int(kick2), int(local_iterations), int(max_full),
                                                    _PRECISIONS[precision])
    if ws_bytes == 0:
        _lib.check("ttamen_amen_solve_workspace_size", 1)
    if prec:
        extra = lib.ttamen_amen_solve_prec_workspace_size(d, Na, rAa, rba, rx0a, int(rmax), int(kickrank),
                                                          int(kick2), prec)
        if extra == 0:
            _lib.check("ttamen_amen_solve_prec_workspace_size", 1)
        ws_bytes += extra
    out_elems = lib.ttamen_amen_solve_output_elems(d, Na, rx0a, int(rmax), int(kickrank), int(kick2))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    xout = torch.empty(max(out_elems, 1), dtype=torch.float64, device=dev)
    rx_out = (ctypes.c_int64 * (d + 1))()
    off_out = (ctypes.c_int64 * d)()
    stats = _lib.TTAmenStats()
    code = lib.ttamen_amen_solve(
        d, Na, rAa, rba, _lib.ptr_array(Ac), _lib.ptr_array(bc),
        _lib.ptr_array(xc) if xc is not None else None, rx0a,
        _lib.ptr_array(z_cores) if z_cores is not None else None,
        int(seed) if seed is not None else 0, int(nswp), float(eps), int(rmax), int(max_full),
        int(kickrank), int(kick2), int(local_iterations), int(resets), int(bool(verbose)), prec,
        _PRECISIONS[precision], ctypes.c_void_p(ws.data_ptr()), ws_bytes,
        ctypes.c_void_p(xout.data_ptr()), out_elems, rx_out, off_out, ctypes.byref(stats),
        ctypes.c_void_p(_lib.stream_handle(dev)))
    _lib.check("ttamen_amen_solve", code)
    cores = []
    for k in range(d):
        n = rx_out[k] * N[k] * rx_out[k + 1]
        c = xout[off_out[k]:off_out[k] + n].reshape(rx_out[k], N[k], rx_out[k + 1]).clone()
        cores.append(c.to(device=out_device, dtype=out_dtype))
    last_stats = stats.as_dict()
    del ws, xout
    return cores, ws_bytes


def amen_divide(a, b, nswp=22, x0=None, eps=1e-10, rmax=100, max_full=500, kickrank=4, kick2=0,
                trunc_norm='res', local_iterations=40, resets=2, verbose=True, preconditioner=None):
    """Elementwise division x = b ./ a of TT tensors by AMEn on the GPU (reference
    torchtt/_division.py:76-398, amen_divide).

    Args follow the reference: `rmax` is an int or a list of d+1 rank caps; x0 defaults to the
    all-ones TT of ranks 1, 2, ..., 2, 1; each GMRES restart runs up to local_iterations + 1
    steps; `preconditioner` is None or 'c'; only `trunc_norm='res'` exists (the reference's
    'fro' branch leaves the rank undefined).

    Returns:
        list[torch.Tensor]: the solution cores, on the device and dtype of a.cores[0].
    """
    global last_stats
    if not (isinstance(a, TT) and isinstance(b, TT)) or a.is_ttm or b.is_ttm:
        raise InvalidArguments('a and b must be TT tensors (not TT-matrices).')
    if list(a.N) != list(b.N):
        raise ShapeMismatch('Both operands should have the same shape.')
    if trunc_norm != 'res':
        raise InvalidArguments("Only trunc_norm='res' is supported.")
    if preconditioner is None:
        prec = 0
    elif preconditioner == 'c':
        prec = 1
    else:
        raise InvalidArguments('Preconditioner ' + str(preconditioner) + ' not defined.')
    _require_real(a, b, x0)
    lib = _lib.load()
    out_device, out_dtype = a.cores[0].device, a.cores[0].dtype
    dev = _solver_device(a)
    N = list(b.N)
    d = len(N)
    caps = [1] + (d - 1) * [int(rmax)] + [1] if isinstance(rmax, int) else [int(v) for v in rmax]
    if len(caps) != d + 1:
        raise InvalidArguments('rmax must be an int or a list of d+1 ranks.')
    if x0 is None:
        rx0 = [1] + (d - 1) * [2] + [1]
        x0c = [torch.ones((rx0[k], N[k], rx0[k + 1]), dtype=out_dtype, device=out_device) for k in range(d)]
    else:
        rx0, x0c = list(x0.R), list(x0.cores)
    # residual-proxy draws after the initial guess, with the options of a (_division.py:98-100)
    z_cores = _draw_z(N, kickrank, kick2, out_device)
    with torch.cuda.device(dev):
        ac = [c.to(device=dev, dtype=torch.float64).contiguous() for c in a.cores]
        bc = [c.to(device=dev, dtype=torch.float64).contiguous() for c in b.cores]
        xc = [c.to(device=dev, dtype=torch.float64).contiguous() for c in x0c]
        if z_cores is not None:
            z_cores = [z.to(device=dev, dtype=torch.float64).contiguous() for z in z_cores]
        Na, raa, rba = _lib.i64_array(N), _lib.i64_array(a.R), _lib.i64_array(b.R)
        rx0a, capsa = _lib.i64_array(rx0), _lib.i64_array(caps)
        ws_bytes = lib.ttamen_amen_divide_workspace_size(d, Na, raa, rba, rx0a, capsa, int(kickrank), int(kick2),
                                                         int(local_iterations), int(max_full), prec)
        if ws_bytes == 0:
            _lib.check("ttamen_amen_divide_workspace_size", 1)
        out_elems = lib.ttamen_amen_divide_output_elems(d, Na, rx0a, capsa, int(kickrank), int(kick2))
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        xout = torch.empty(max(out_elems, 1), dtype=torch.float64, device=dev)
        rx_out = (ctypes.c_int64 * (d + 1))()
        off_out = (ctypes.c_int64 * d)()
        stats = _lib.TTAmenStats()
        code = lib.ttamen_amen_divide(
            d, Na, raa, rba, _lib.ptr_array(ac), _lib.ptr_array(bc), _lib.ptr_array(xc), rx0a,
            _lib.ptr_array(z_cores) if z_cores is not None else None, int(nswp), float(eps), capsa,
            int(max_full), int(kickrank), int(kick2), int(local_iterations), int(resets), int(bool(verbose)),
            prec, ctypes.c_void_p(ws.data_ptr()), ws_bytes, ctypes.c_void_p(xout.data_ptr()), out_elems,
            rx_out, off_out, ctypes.byref(stats), ctypes.c_void_p(_lib.stream_handle(dev)))
        _lib.check("ttamen_amen_divide", code)
        cores = []
        for k in range(d):
            n = rx_out[k] * N[k] * rx_out[k + 1]
            c = xout[off_out[k]:off_out[k] + n].reshape(rx_out[k], N[k], rx_out[k + 1]).clone()
            cores.append(c.to(device=out_device, dtype=out_dtype))
        last_stats = stats.as_dict()
        last_stats["workspace_bytes"] = int(ws_bytes)
        del ws, xout
    return cores


#: per-system statistics of the most recent batch (list of dicts)
last_batch_stats = None


def _build_systems(systems, dev, kickrank, kick2, z0, lib):
    """ctypes array of ttamen_system for (A, b[, rmax[, x0]]) items, with device copies and outputs."""
    count = len(systems)
    keep, outs = [], []
    arr = (_lib.TTAmenSystem * count)()
    stats = (_lib.TTAmenStats * count)()
    for i, item in enumerate(systems):
        A, b = item[0], item[1]
        rmax = int(item[2]) if len(item) > 2 and item[2] is not None else 32768
        x0 = item[3] if len(item) > 3 else None
        _check_system(A, b)
        _require_real(x0)
        N = list(b.N)
        d = len(N)
        Ac = [c.to(device=dev, dtype=torch.float64).contiguous() for c in A.cores]
        bc = [c.to(device=dev, dtype=torch.float64).contiguous() for c in b.cores]
        if x0 is not None:
            if list(x0.N) != N:
                raise ShapeMismatch('x0 does not match the shape of b.')
            xc = [c.to(device=dev, dtype=torch.float64).contiguous() for c in x0.cores]
            rx0a, px0 = _lib.i64_array(x0.R), _lib.ptr_array(xc)
        else:
            xc, rx0a, px0 = None, None, None
        zc = z0[i] if z0 is not None else _draw_z(N, kickrank, kick2, A.cores[0].device)
        if zc is not None:
            zc = [z.to(device=dev, dtype=torch.float64).contiguous() for z in zc]
        Na, rAa, rba = _lib.i64_array(N), _lib.i64_array(A.R), _lib.i64_array(b.R)
        out_elems = lib.ttamen_amen_solve_output_elems(d, Na, rx0a, rmax, int(kickrank), int(kick2))
        xout = torch.empty(max(out_elems, 1), dtype=torch.float64, device=dev)
        rx_out = (ctypes.c_int64 * (d + 1))()
        off_out = (ctypes.c_int64 * d)()
        pa, pb = _lib.ptr_array(Ac), _lib.ptr_array(bc)
        pz = _lib.ptr_array(zc) if zc is not None else None
        keep.extend([Ac, bc, zc, xc, rx0a, px0, Na, rAa, rba, pa, pb, pz, xout, rx_out, off_out])
        y = arr[i]
        y.d = d
        y.N, y.rA, y.rb = Na, rAa, rba
        y.A_cores, y.b_cores = pa, pb
        y.x0_cores, y.rx0 = px0, rx0a
        y.z0_cores = pz
        y.seed = 0
        y.rmax = rmax
        y.x_out = ctypes.c_void_p(xout.data_ptr())
        y.x_out_elems = out_elems
        y.rx_out, y.offsets_out = rx_out, off_out
        y.stats = ctypes.pointer(stats[i])
        y.status = 0
        outs.append((A, N, xout, rx_out, off_out))
    return arr, stats, outs, keep


def _collect(outs, statuses=None):
    result = []
    for i, (A, N, xout, rx_out, off_out) in enumerate(outs):
        if statuses is not None and statuses[i] != 0:
            result.append(None)
            continue
        d = len(N)
        cores = []
        for k in range(d):
            n = rx_out[k] * N[k] * rx_out[k + 1]
            c = xout[off_out[k]:off_out[k] + n].reshape(rx_out[k], N[k], rx_out[k + 1]).clone()
            cores.append(c.to(device=A.cores[0].device, dtype=A.cores[0].dtype))
        result.append(TT(cores))
    return result


#: per-system statistics of the most recent batch (list of dicts)
last_batch_stats = None


def _check_batch(fn, code, arr, outs):
    """Raise BatchError (every system's status, the solved systems' results) when `code` != 0."""
    if code == 0:
        return
    statuses = [int(arr[i].status) for i in range(len(outs))]
    msg = _lib.load().ttamen_last_error()
    raise _lib.BatchError(fn, code, msg.decode() if msg else "", statuses, _collect(outs, statuses))


def amen_solve_batch(systems, nswp=22, eps=1e-10, max_full=500, kickrank=4, kick2=0, local_iterations=40,
                     resets=2, precision='fp64', workers=8, z0=None):
    """Solve many independent TT systems concurrently (no counterpart in the reference).

    systems: list of (A, b), (A, b, rmax) or (A, b, rmax, x0) with TT A, b (and warm start x0).
    Each system runs the same solver as
    `amen_solve` with the same options; the systems run together as one cohort
    (libttamen: ttamen_amen_solve_batch): at every host synchronisation point the kernels
    that all running systems have queued are merged into grouped launches. `workers` host
    threads run the systems' host code. The residual-proxy draws are made per system in list
    order (or taken from `z0`, a list of per-system core lists), so system i's result equals
    `amen_solve(A_i, b_i, ..., _z0=z0[i])` bit for bit.

    Returns:
        list[TT]: the solutions, on each A's device and dtype.

    Raises:
        BatchError: when a system fails; it carries every system's status (`statuses`) and the
        solutions of the systems that succeeded (`results`, None where a system failed). Its
        code is the first failing system's status.
    """
    global last_batch_stats
    if precision not in _PRECISIONS:
        raise InvalidArguments("precision must be one of %s." % sorted(_PRECISIONS))
    lib = _lib.load()
    count = len(systems)
    if count == 0:
        return []
    dev = _solver_device(systems[0][0])
    with torch.cuda.device(dev):
        arr, stats, outs, keep = _build_systems(systems, dev, kickrank, kick2, z0, lib)
        workers = max(1, min(int(workers), count))
        ws_bytes = lib.ttamen_amen_solve_batch_workspace_size(arr, count, int(kickrank), int(kick2),
                                                              int(local_iterations), int(max_full),
                                                              _PRECISIONS[precision], workers)
        if ws_bytes == 0:
            _lib.check("ttamen_amen_solve_batch_workspace_size", 1)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        code = lib.ttamen_amen_solve_batch(arr, count, int(nswp), float(eps), int(max_full), int(kickrank),
                                           int(kick2), int(local_iterations), int(resets),
                                           _PRECISIONS[precision], workers, ctypes.c_void_p(ws.data_ptr()),
                                           ws_bytes, ctypes.c_void_p(_lib.stream_handle(dev)))
        last_batch_stats = [stats[i].as_dict() for i in range(count)]
        _check_batch("ttamen_amen_solve_batch", code, arr, outs)
        result = _collect(outs)
        cohort_stats()
        del ws, keep
    return result


def gemm_autotune(mode=None):
    """GEMM tile autotuner of single-system solves: mode 1 times the four tile shapes on the first
    launch of each shape bucket and keeps the fastest, mode 0 (default) uses the fixed rule; None
    only queries. Returns the previous mode."""
    lib = _lib.load()
    return int(lib.ttamen_gemm_autotune(-1 if mode is None else int(mode)))


def gemm_autotune_stats():
    """Buckets tuned, launches served from the cache, seconds spent tuning, launches off the rule."""
    import ctypes
    lib = _lib.load()
    out = (ctypes.c_double * 4)()
    _lib.check("ttamen_gemm_autotune_stats", lib.ttamen_gemm_autotune_stats(out))
    return {"buckets_tuned": int(out[0]), "launches_from_cache": int(out[1]), "tuning_s": out[2],
            "launches_off_rule": int(out[3])}


def round_tt(x, eps=1e-12, rmax=sys.maxsize):
    """TT rounding on the GPU (libttamen: ttamen_tt_round), replacing the reference's native
    `round_this` (cpp/compression.h:3-41): left-to-right QR, then right-to-left truncated SVD
    with accuracy eps/sqrt(d-1) per bond (rank_chop, cpp/ortho.h:16-44) and rank cap rmax.
    FP64 cores on one device; ranks up to 140. Returns a new TT."""
    if not isinstance(x, TT):
        raise InvalidArguments("Input must be a TT.")
    _require_real(x)
    lib = _lib.load()
    dev = x.cores[0].device
    if dev.type != "cuda":
        raise RuntimeError("round_tt runs on a ROCm GPU; the input cores are on %s" % dev)
    d = len(x.cores)
    shapes = [tuple(c.shape[1:-1]) for c in x.cores]
    N = [int(np.prod(sh)) for sh in shapes]
    R = list(x.R)
    cores = [c.to(torch.float64).contiguous() for c in x.cores]
    Na, Ra = _lib.i64_array(N), _lib.i64_array(R)
    with torch.cuda.device(dev):
        ws_bytes = lib.ttamen_tt_round_workspace_size(d, Na, Ra)
        if ws_bytes == 0:
            _lib.check("ttamen_tt_round_workspace_size", 1)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        total = sum(R[k] * N[k] * R[k + 1] for k in range(d))
        out = torch.empty(total, dtype=torch.float64, device=dev)
        r_out = (ctypes.c_int64 * (d + 1))()
        off = (ctypes.c_int64 * d)()
        rmax = int(min(rmax, 2 ** 62))
        code = lib.ttamen_tt_round(d, Na, Ra, _lib.ptr_array(cores), float(eps), rmax,
                                   ctypes.c_void_p(ws.data_ptr()), ws_bytes, ctypes.c_void_p(out.data_ptr()),
                                   total, r_out, off, ctypes.c_void_p(_lib.stream_handle(dev)))
        _lib.check("ttamen_tt_round", code)
        res = [out[off[k]:off[k] + r_out[k] * N[k] * r_out[k + 1]].reshape((r_out[k],) + shapes[k] + (r_out[k + 1],))
               .clone().to(x.cores[0].dtype) for k in range(d)]
        del ws
    return TT(res)


def tt_full(x):
    """Full tensor of a TT vector on the GPU (libttamen: ttamen_tt_full), replacing the reference's
    native `tt_full` (cpp/full.h:3-12): the cores multiplied left to right, one GEMM per core.
    FP64 cores on one device; returns an N[0] x ... x N[d-1] tensor."""
    if not isinstance(x, TT):
        raise InvalidArguments("Input must be a TT.")
    if x.is_ttm:
        raise IncompatibleTypes("tt_full expects a TT vector (the reference's native full is for 3-D cores).")
    _require_real(x)
    lib = _lib.load()
    dev = x.cores[0].device
    if dev.type != "cuda":
        raise RuntimeError("tt_full runs on a ROCm GPU; the input cores are on %s" % dev)
    d = len(x.cores)
    N, R = list(x.N), list(x.R)
    cores = [c.to(torch.float64).contiguous() for c in x.cores]
    Na, Ra = _lib.i64_array(N), _lib.i64_array(R)
    total = int(np.prod(N))
    with torch.cuda.device(dev):
        ws_bytes = lib.ttamen_tt_full_workspace_size(d, Na, Ra)
        if ws_bytes == 0:
            _lib.check("ttamen_tt_full_workspace_size", 1)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        out = torch.empty(total, dtype=torch.float64, device=dev)
        code = lib.ttamen_tt_full(d, Na, Ra, _lib.ptr_array(cores), ctypes.c_void_p(ws.data_ptr()), ws_bytes,
                                  ctypes.c_void_p(out.data_ptr()), total, ctypes.c_void_p(_lib.stream_handle(dev)))
        _lib.check("ttamen_tt_full", code)
        del ws
    return out.reshape(N).to(x.cores[0].dtype)


#: counters of the most recent batched / served call (see cohort_stats)
last_cohort_stats = None


def cohort_stats():
    """Counters of the last batched or served call: rounds, grouped launches, merged launch
    items, member resumptions and the seconds spent resuming members, building launches,
    waiting for the device and in total (libttamen: ttamen_cohort_stats)."""
    global last_cohort_stats
    lib = _lib.load()
    out = (ctypes.c_double * 8)()
    _lib.check("ttamen_cohort_stats", lib.ttamen_cohort_stats(out))
    keys = ("rounds", "launches", "items", "resumes", "t_resume", "t_build", "t_gpu", "t_total")
    last_cohort_stats = dict(zip(keys, list(out)))
    return last_cohort_stats


def residual_norms(A, x, b):
    """(||A x - b||, ||b||) in FP64 on the GPU (libttamen: ttamen_tt_residual_norm): a left-to-right
    sweep over the residual TT (ranks rA rx + rb) with Householder R factors, replacing the TT algebra
    the reference's callers use for this check ((A @ x - b).norm(), torchtt/_tt_base.py:780-846,
    :545, :990-1052). Inputs on the CPU are copied to the GPU."""
    if not (isinstance(A, TT) and isinstance(x, TT) and isinstance(b, TT)):
        raise InvalidArguments('A, x and b must be TT instances.')
    if not A.is_ttm or x.is_ttm or b.is_ttm:
        raise IncompatibleTypes('A must be TT-matrix, x and b vectors.')
    if list(A.N) != list(x.N) or list(A.M) != list(b.N):
        raise ShapeMismatch('Dimension mismatch.')
    _require_real(A, x, b)
    lib = _lib.load()
    dev = _solver_device(A)
    N = list(b.N)
    d = len(N)
    with torch.cuda.device(dev):
        Ac = [c.to(device=dev, dtype=torch.float64).contiguous() for c in A.cores]
        xc = [c.to(device=dev, dtype=torch.float64).contiguous() for c in x.cores]
        bc = [c.to(device=dev, dtype=torch.float64).contiguous() for c in b.cores]
        Na, rA, rx, rb = (_lib.i64_array(v) for v in (N, A.R, x.R, b.R))
        wsb = lib.ttamen_tt_residual_norm_workspace_size(d, Na, rA, rx, rb)
        if wsb == 0:
            _lib.check("ttamen_tt_residual_norm_workspace_size", 1)
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        out = (ctypes.c_double * 2)()
        code = lib.ttamen_tt_residual_norm(d, Na, rA, rx, rb, _lib.ptr_array(Ac), _lib.ptr_array(xc),
                                           _lib.ptr_array(bc), ctypes.c_void_p(ws.data_ptr()), wsb, out,
                                           ctypes.c_void_p(_lib.stream_handle(dev)))
        _lib.check("ttamen_tt_residual_norm", code)
        del ws
    return float(out[0]), float(out[1])


def residual(A, x, b):
    """Relative residual ||A x - b|| / ||b|| in FP64 on the GPU (the check every test applies);
    see residual_norms."""
    r, nb = residual_norms(A, x, b)
    return r / nb
