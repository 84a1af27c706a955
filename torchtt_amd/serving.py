"""Serving loop for AMEn solve requests (continuous batching; no counterpart in the reference).

`serve(requests, arrivals, ...)` hands a whole request stream to libttamen's native loop
(include/ttamen.h: ttamen_serve): requests are admitted in arrival order whenever a slot is
idle and a contiguous region of their footprint (the solver workspace at their rank cap) is
free in the serving memory. The running requests form one cohort: the same sweep steps of
different requests run as one grouped launch, with each request's cores and interfaces
resident in its region. Results are the same as solving each request alone.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .solvers import _PRECISIONS, _build_systems, _check_batch, _collect, _solver_device
from .errors import InvalidArguments


def poisson_arrivals(rate, duration, seed=0):
    """Arrival times (s) of a Poisson process of `rate` requests/s over `duration` seconds."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append(t)


def serve(requests, arrivals, slots=64, memory_bytes=None, nswp=22, eps=1e-10, max_full=500, kickrank=4,
          kick2=0, local_iterations=40, resets=2, precision='fp64', z0=None):
    """Serve `requests` ((A, b[, rmax[, x0]]) items) arriving at `arrivals` (seconds, sorted).

    slots: requests running at once (at most 256).
    memory_bytes: serving memory (HBM) budget shared by the running requests; default
    slots x the largest request footprint. The launch-table staging area
    (ttamen_cohort_staging_bytes) is allocated on top of it.

    Returns:
        (solutions, info) where info holds per-request arrival/start/done times (s),
        latency = done - arrival, and the per-request solver statistics.
    """
    if precision not in _PRECISIONS:
        raise InvalidArguments("precision must be one of %s." % sorted(_PRECISIONS))
    count = len(requests)
    if count != len(arrivals):
        raise InvalidArguments("one arrival time per request")
    if count == 0:
        return [], {}
    lib = _lib.load()
    dev = _solver_device(requests[0][0])
    with torch.cuda.device(dev):
        arr, stats, outs, keep = _build_systems(requests, dev, kickrank, kick2, z0, lib)
        if memory_bytes is None:
            need = 0
            for i in range(count):
                y = arr[i]
                w = lib.ttamen_amen_solve_workspace_size(y.d, y.N, y.rA, y.rb, y.rx0, y.rmax, int(kickrank),
                                                         int(kick2), int(local_iterations), int(max_full),
                                                         _PRECISIONS[precision])
                if w == 0:
                    _lib.check("ttamen_amen_solve_workspace_size", 1)
                need = max(need, w)
            memory_bytes = min(int(slots), count) * (need + 256)
        ws_bytes = int(memory_bytes) + int(lib.ttamen_cohort_staging_bytes())
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        arr_t = (ctypes.c_double * count)(*[float(a) for a in arrivals])
        start = (ctypes.c_double * count)()
        done = (ctypes.c_double * count)()
        code = lib.ttamen_serve(arr, arr_t, count, int(nswp), float(eps), int(max_full), int(kickrank), int(kick2),
                                int(local_iterations), int(resets), _PRECISIONS[precision], int(slots),
                                ctypes.c_void_p(ws.data_ptr()), ws_bytes, start, done,
                                ctypes.c_void_p(_lib.stream_handle(dev)))
        _check_batch("ttamen_serve", code, arr, outs)
        sols = _collect(outs)
        arrival = np.array(list(arrivals), dtype=float)
        info = {
            "arrival_s": arrival,
            "start_s": np.array(start[:]),
            "done_s": np.array(done[:]),
            "latency_s": np.array(done[:]) - arrival,
            "stats": [stats[i].as_dict() for i in range(count)],
            "memory_bytes": int(memory_bytes),
        }
        del ws, keep
    return sols, info
