"""DMRG approximation of a TT matrix-vector product on the GPU, with the reference's signature
(torchtt/_dmrg.py:22-45, TT.fast_matvec at torchtt/_tt_base.py:847-873). The sweep runs in
libttamen (ttamen_dmrg_mv, the restatement of cpp/dmrg_mv.h); there is no CPU path.
"""
import ctypes

import torch

from ._tt import TT
from .errors import InvalidArguments, IncompatibleTypes, ShapeMismatch
from . import _lib
from .solvers import _solver_device

#: statistics of the most recent call (dict)
last_stats = None


def dmrg_matvec(A, x, y0=None, nswp=20, eps=1e-12, rmax=32768, kickrank=4, verb=False, use_cpp=True, _seed=None):
    """y ~= A @ x by DMRG sweeps (reference torchtt/_dmrg.py:22-45 -> cpp/dmrg_mv.h:23-230).

    Args follow the reference: `y0` an initial guess (None: random cores of rank 2 drawn with
    torch.randn, as the reference's native path does), `rmax` an int or a list of d+1 ranks,
    `use_cpp` accepted for compatibility (the native sweep always runs). Real FP64 arithmetic;
    at least two cores. `_seed` fixes the device stream of the kick columns (default: drawn
    from torch's generator, so torch.manual_seed makes a call repeatable).

    Returns:
        TT: the product, on the device and dtype of A.cores[0].
    """
    global last_stats
    if not (isinstance(A, TT) and isinstance(x, TT)):
        raise InvalidArguments('Second operand has to be TT object.')
    if not A.is_ttm or x.is_ttm:
        raise IncompatibleTypes('First operand should be a TT matrix and second a TT vector.')
    if list(A.N) != list(x.N):
        raise ShapeMismatch('Dimension mismatch.')
    if A.cores[0].is_complex() or x.cores[0].is_complex():
        raise InvalidArguments('dmrg_matvec computes in real FP64; complex cores are not supported.')
    d = len(A.N)
    if d < 2:
        raise InvalidArguments('dmrg_matvec needs at least two cores.')
    M, N = list(A.M), list(A.N)
    caps = [1] + (d - 1) * [int(rmax)] + [1] if isinstance(rmax, int) else [int(v) for v in rmax]
    if len(caps) != d + 1:
        raise InvalidArguments('rmax must be an int or a list of d+1 ranks.')
    dev = _solver_device(A)
    out_device, out_dtype = A.cores[0].device, A.cores[0].dtype
    lib = _lib.load()
    with torch.cuda.device(dev):
        Ac = [c.to(device=dev, dtype=torch.float64).contiguous() for c in A.cores]
        xc = [c.to(device=dev, dtype=torch.float64).contiguous() for c in x.cores]
        if y0 is None:
            ry0 = [1] + (d - 1) * [2] + [1]
            yc = [torch.randn((ry0[k], M[k], ry0[k + 1]), dtype=torch.float64, device=dev) for k in range(d)]
        else:
            if y0.is_ttm or list(y0.N) != M:
                raise ShapeMismatch('The initial guess must be a TT vector of the shape of A @ x.')
            ry0 = list(y0.R)
            yc = [c.to(device=dev, dtype=torch.float64).contiguous() for c in y0.cores]
        seed = int(torch.randint(0, 2 ** 62, (1,)).item()) if _seed is None else int(_seed)
        Ma, Na, rAa, rxa = _lib.i64_array(M), _lib.i64_array(N), _lib.i64_array(A.R), _lib.i64_array(x.R)
        ry0a, capsa = _lib.i64_array(ry0), _lib.i64_array(caps)
        ws_bytes = lib.ttamen_dmrg_mv_workspace_size(d, Ma, Na, rAa, rxa, ry0a, capsa, int(kickrank))
        out_elems = lib.ttamen_dmrg_mv_output_elems(d, Ma, ry0a, capsa, int(kickrank))
        if ws_bytes == 0 or out_elems == 0:
            _lib.check("ttamen_dmrg_mv_workspace_size", 1)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        yout = torch.empty(out_elems, dtype=torch.float64, device=dev)
        ry_out = (ctypes.c_int64 * (d + 1))()
        off_out = (ctypes.c_int64 * d)()
        stats = _lib.TTAmenStats()
        code = lib.ttamen_dmrg_mv(d, Ma, Na, rAa, rxa, _lib.ptr_array(Ac), _lib.ptr_array(xc), _lib.ptr_array(yc),
                                  ry0a, int(nswp), float(eps), capsa, int(kickrank), seed, int(bool(verb)),
                                  ctypes.c_void_p(ws.data_ptr()), ws_bytes, ctypes.c_void_p(yout.data_ptr()),
                                  out_elems, ry_out, off_out, ctypes.byref(stats),
                                  ctypes.c_void_p(_lib.stream_handle(dev)))
        _lib.check("ttamen_dmrg_mv", code)
        cores = []
        for k in range(d):
            n = ry_out[k] * M[k] * ry_out[k + 1]
            cores.append(yout[off_out[k]:off_out[k] + n].reshape(ry_out[k], M[k], ry_out[k + 1]).clone()
                         .to(device=out_device, dtype=out_dtype))
        last_stats = stats.as_dict()
        del ws, yout
    return TT(cores)
