"""AMEn approximation of TT products on the GPU: amen_mv (TT-matrix times TT vector) and amen_mm
(TT-matrix times TT-matrix), with the reference's signatures (torchtt/_amen.py:55-123). The sweep
runs in libttamen (ttamen_amen_mv / ttamen_amen_mm); there is no CPU path.
"""
import ctypes

import torch

from ._tt import TT
from .errors import InvalidArguments, IncompatibleTypes, ShapeMismatch
from . import _lib
from .solvers import _solver_device, _require_real

#: statistics of the most recent call (dict)
last_stats = None


def _caps(rmax, d):
    caps = [1] + (d - 1) * [int(rmax)] + [1] if isinstance(rmax, int) else [int(v) for v in rmax]
    if len(caps) != d + 1:
        raise InvalidArguments('rmax must be an int or a list of d+1 ranks.')
    return caps


def _draw_z(P, kickrank, kick2, device):
    """Residual-proxy cores drawn as the reference's torchtt.random(..., rz) (_amen.py:146-151):
    N(0,1) cores of ranks 1, rz, ..., rz, 1, one after another."""
    d = len(P)
    rz = [1] + (d - 1) * [kickrank + kick2] + [1]
    if kickrank + kick2 <= 0:
        return None
    return [torch.randn((rz[k], P[k], rz[k + 1]), dtype=torch.float64, device=device) for k in range(d)]


def _run(entry, sizes, A_cores, b_cores, P, x0, kickrank, kick2, rmax, nswp, eps, verbose, dev, device, dtype):
    lib = _lib.load()
    d = len(P)
    caps = _caps(rmax, d)
    z = _draw_z(P, kickrank, kick2, device)
    with torch.cuda.device(dev):
        Ac = [c.to(device=dev, dtype=torch.float64).contiguous() for c in A_cores]
        bc = [c.to(device=dev, dtype=torch.float64).contiguous() for c in b_cores]
        if x0 is not None:
            xc = [c.to(device=dev, dtype=torch.float64).reshape(c.shape[0], -1, c.shape[-1]).contiguous() for c in x0.cores]
            rx0 = list(x0.R)
        else:
            xc, rx0 = None, None
        if z is not None:
            z = [c.to(device=dev, dtype=torch.float64).contiguous() for c in z]
        capsa = _lib.i64_array(caps)
        rx0a = _lib.i64_array(rx0) if rx0 is not None else None
        cap = None
        while True:  # at the default rank capacity, then larger ones if the sweep outgrows it
            prev_cap = lib.ttamen_rank_capacity(cap) if cap else None
            try:
                cores = _run_native(entry, sizes, lib, d, P, Ac, bc, xc, rx0a, z, capsa, kickrank, kick2, nswp, eps,
                                    verbose, dev, device, dtype)
                break
            except _lib.TTAmenError as e:
                from .solvers import _grown_capacity
                cap = _grown_capacity(e, cap or lib.ttamen_rank_capacity(0))
                if cap is None:
                    raise
            finally:
                if prev_cap is not None:
                    lib.ttamen_rank_capacity(prev_cap)
    return cores


def _run_native(entry, sizes, lib, d, P, Ac, bc, xc, rx0a, z, capsa, kickrank, kick2, nswp, eps, verbose, dev,
                device, dtype):
    """One call of the native product at the calling thread's rank capacity; the product's cores."""
    global last_stats
    ws_bytes, out_elems, call = sizes(lib, capsa, rx0a)
    if ws_bytes == 0 or out_elems == 0:
        _lib.check(entry + "_workspace_size", 1)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    xout = torch.empty(out_elems, dtype=torch.float64, device=dev)
    rx_out = (ctypes.c_int64 * (d + 1))()
    off_out = (ctypes.c_int64 * d)()
    stats = _lib.TTAmenStats()
    code = call(lib, _lib.ptr_array(Ac), _lib.ptr_array(bc), _lib.ptr_array(xc) if xc is not None else None, rx0a,
                _lib.ptr_array(z) if z is not None else None, int(nswp), float(eps), capsa, int(kickrank),
                int(kick2), 0, int(bool(verbose)), ctypes.c_void_p(ws.data_ptr()), ws_bytes,
                ctypes.c_void_p(xout.data_ptr()), out_elems, rx_out, off_out, ctypes.byref(stats),
                ctypes.c_void_p(_lib.stream_handle(dev)))
    _lib.check(entry, code)
    cores = []
    for k in range(d):
        n = rx_out[k] * P[k] * rx_out[k + 1]
        cores.append(xout[off_out[k]:off_out[k] + n].reshape(rx_out[k], P[k], rx_out[k + 1]).clone()
                     .to(device=device, dtype=dtype))
    last_stats = stats.as_dict()
    del ws, xout
    return cores


def amen_mv(A, b, nswp=22, x0=None, eps=1e-10, rmax=1024, kickrank=4, kick2=0, verbose=False, use_cpp=True):
    """y ~= A @ b by AMEn (reference torchtt/_amen.py:55-101), suited to a low-rank product.

    Args follow the reference (`use_cpp` is accepted for compatibility: the reference always
    takes its Python path, which this sweep restates on the GPU).

    Returns:
        TT: the approximation, on the device and dtype of A.cores[0].
    """
    if not (isinstance(A, TT) and isinstance(b, TT)):
        raise InvalidArguments('A and b must be TT instances.')
    if not (A.is_ttm and not b.is_ttm):
        raise IncompatibleTypes('A must be TT-matrix and b must be vector.')
    if list(A.N) != list(b.N):
        raise ShapeMismatch('Dimension mismatch.')
    _require_real(A, b, x0)
    d = len(A.M)
    M, K = list(A.M), list(A.N)
    Ma, Ka, rAa, rba = _lib.i64_array(M), _lib.i64_array(K), _lib.i64_array(A.R), _lib.i64_array(b.R)

    def sizes(lib, capsa, rx0a):
        ws = lib.ttamen_amen_mv_workspace_size(d, Ma, Ka, rAa, rba, rx0a, capsa, int(kickrank), int(kick2))
        out = lib.ttamen_amen_mv_output_elems(d, Ma, Ka, rx0a, capsa, int(kickrank), int(kick2))

        def call(lib, *args):
            return lib.ttamen_amen_mv(d, Ma, Ka, rAa, rba, *args)
        return ws, out, call

    cores = _run("ttamen_amen_mv", sizes, A.cores, b.cores, M, x0, kickrank, kick2, rmax, nswp, eps, verbose,
                 _solver_device(A), A.cores[0].device, A.cores[0].dtype)
    return TT(cores)


def amen_mm(A, B, nswp=22, X0=None, eps=1e-10, rmax=1024, kickrank=4, kick2=0, verbose=False):
    """C ~= A @ B for TT-matrices by AMEn (reference torchtt/_amen.py:104-123).

    Returns:
        TT: the TT-matrix approximation, on the device and dtype of A.cores[0].
    """
    if not (isinstance(A, TT) and isinstance(B, TT)) or not (A.is_ttm and B.is_ttm):
        raise InvalidArguments('A and B must be TT-matrices.')
    if list(A.N) != list(B.M):
        raise ShapeMismatch('Dimension mismatch.')
    _require_real(A, B, X0)
    d = len(A.M)
    M, K, N = list(A.M), list(A.N), list(B.N)
    P = [m * n for m, n in zip(M, N)]
    Ma, Ka, Na = _lib.i64_array(M), _lib.i64_array(K), _lib.i64_array(N)
    rAa, rBa = _lib.i64_array(A.R), _lib.i64_array(B.R)
    Pa, KNa = _lib.i64_array(P), _lib.i64_array([k * n for k, n in zip(K, N)])

    def sizes(lib, capsa, rx0a):
        ws = lib.ttamen_amen_mm_workspace_size(d, Ma, Ka, Na, rAa, rBa, rx0a, capsa, int(kickrank), int(kick2))
        out = lib.ttamen_amen_mv_output_elems(d, Pa, KNa, rx0a, capsa, int(kickrank), int(kick2))

        def call(lib, *args):
            return lib.ttamen_amen_mm(d, Ma, Ka, Na, rAa, rBa, *args)
        return ws, out, call

    B_vec = [c.reshape(c.shape[0], -1, c.shape[-1]) for c in B.cores]
    cores = _run("ttamen_amen_mm", sizes, A.cores, B_vec, P, X0, kickrank, kick2, rmax, nswp, eps, verbose,
                 _solver_device(A), A.cores[0].device, A.cores[0].dtype)
    return TT([c.reshape(c.shape[0], m, n, c.shape[-1]) for c, m, n in zip(cores, M, N)])
