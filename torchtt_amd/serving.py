"""Serving loop for AMEn solve requests (continuous batching; no counterpart in the reference).

`Server` (include/ttamen.h: ttamen_server_*) takes requests one at a time while others run.
`serve(requests, arrivals, ...)` replays a whole request stream through libttamen's native loop
(ttamen_serve): requests are admitted in arrival order whenever a slot is
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


class Request:
    """A request submitted to a Server: poll it with done(), collect it with result()."""

    def __init__(self, server, rid, keep, out, stats):
        self._server, self.id, self._keep, self._out, self._stats = server, rid, keep, out, stats
        self._result = None

    def state(self):
        """0 queued, 1 running, 2 done."""
        st = ctypes.c_int()
        _lib.check("ttamen_server_poll", self._server._lib.ttamen_server_poll(self._server._h, self.id,
                                                                             ctypes.byref(st), None))
        return st.value

    def done(self):
        return self.state() == 2

    def times(self):
        """(submitted, admitted, completed) in seconds since the server started."""
        t = (ctypes.c_double * 3)()
        _lib.check("ttamen_server_poll", self._server._lib.ttamen_server_poll(self._server._h, self.id, None, t))
        return tuple(t)

    def result(self):
        """Block until the request is done and return its solution (a TT on the device and dtype of A)."""
        if self._result is None:
            code = self._server._lib.ttamen_server_wait(self._server._h, self.id)
            _lib.check("ttamen_server_wait", code)
            self._result = _collect([self._out])[0]
            self.stats = self._stats[0].as_dict()
            self._keep = None
        return self._result


class Server:
    """Online serving (libttamen: ttamen_server_*): requests are submitted while others run and are
    admitted, in submission order, whenever a slot is idle and a contiguous region of their footprint
    is free in the serving memory; the running requests share grouped launches (continuous batching).

    Shared options are fixed here; rank cap (and optionally z0 draws) come with each request. Each
    result equals `amen_solve` on the same inputs. Use as a context manager or call close(), which
    lets every submitted request finish.
    """

    def __init__(self, slots=64, memory_bytes=4 << 30, nswp=22, eps=1e-10, max_full=500, kickrank=4, kick2=0,
                 local_iterations=40, resets=2, precision='fp64', device=None):
        if precision not in _PRECISIONS:
            raise InvalidArguments("precision must be one of %s." % sorted(_PRECISIONS))
        self._lib = _lib.load()
        self._dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self._kick = (int(kickrank), int(kick2))
        ws_bytes = int(memory_bytes) + int(self._lib.ttamen_cohort_staging_bytes())
        with torch.cuda.device(self._dev):
            self._ws = torch.empty(ws_bytes, dtype=torch.uint8, device=self._dev)
            h = ctypes.c_void_p()
            code = self._lib.ttamen_server_create(ctypes.byref(h), int(nswp), float(eps), int(max_full), int(kickrank),
                                                  int(kick2), int(local_iterations), int(resets),
                                                  _PRECISIONS[precision], int(slots),
                                                  ctypes.c_void_p(self._ws.data_ptr()), ws_bytes)
            _lib.check("ttamen_server_create", code)
        self._h = h
        self._open = True

    def submit(self, A, b, rmax=32768, z0=None, x0=None):
        """Queue one solve of A x = b; returns a Request at once."""
        if not self._open:
            raise InvalidArguments("the server is closed")
        with torch.cuda.device(self._dev):
            arr, stats, outs, keep = _build_systems([(A, b, rmax, x0)], self._dev, self._kick[0], self._kick[1],
                                                    None if z0 is None else [z0], self._lib)
            rid = ctypes.c_int64()
            code = self._lib.ttamen_server_submit(self._h, arr, ctypes.c_void_p(_lib.stream_handle(self._dev)),
                                                  ctypes.byref(rid))
            _lib.check("ttamen_server_submit", code)
        return Request(self, rid.value, (arr, stats, keep), outs[0], stats)

    def close(self):
        if self._open:
            self._open = False
            _lib.check("ttamen_server_destroy", self._lib.ttamen_server_destroy(self._h))
            self._ws = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
