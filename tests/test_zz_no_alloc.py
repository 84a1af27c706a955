"""The allocation-free local operator (ttamen_local_matvec_noalloc) against the workspace one
(ttamen_local_matvec): the same kernels in the same order on the same inputs, so the results are
bitwise equal; both within 1e-12 of a torch FP64 contraction. Three calls in a row on one GPU
(each later one smaller), then a size above the scratch limit, refused with status 2 before any
device work."""
import ctypes
import sys

import pytest
import torch

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _operands(l1, s, r1, n, S, l2, r2, nvec, seed):
    g = torch.Generator().manual_seed(seed)
    mk = lambda *sh: torch.randn(*sh, dtype=torch.float64, generator=g).cuda()  # noqa: E731
    return mk(l1, s, r1), mk(s, n, n, S), mk(l2, S, r2), mk(nvec, r1, n, r2)


def _with_workspace(lib, _lib, PhiL, A, PhiR, x, dims, stream):
    l1, s, r1, n, S, l2, r2, nvec = dims
    y = torch.full((nvec, l1, n, l2), float("nan"), dtype=torch.float64, device="cuda")
    wsb = lib.ttamen_local_matvec_workspace_size(l1, s, r1, n, S, l2, r2, nvec)
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    _lib.check("ttamen_local_matvec", lib.ttamen_local_matvec(_p(PhiL), _p(A), _p(PhiR), _p(x), _p(y), *dims,
                                                              _p(ws), wsb, ctypes.c_void_p(stream)))
    return y


def test_noalloc_matvec_equals_workspace_matvec():
    from torchtt_amd import _lib
    lib = _lib.load()
    stream = _lib.stream_handle(torch.device("cuda", 0))
    limit = lib.ttamen_local_matvec_noalloc_limit()
    assert limit == 64 << 20
    cases = [
        (2, 4, 6, 160, 4, 3, 8, 1),      # middle stage split over K (K = s n = 640, 20 output tiles)
        (48, 4, 44, 64, 3, 36, 40, 2),   # interface-sized ranks, two vectors
        (8, 2, 10, 12, 2, 6, 5, 1),      # small
    ]
    for i, dims in enumerate(cases):
        l1, s, r1, n, S, l2, r2, nvec = dims
        _say("noalloc call %d: l1=%d s=%d r1=%d n=%d S=%d l2=%d r2=%d nvec=%d" % ((i,) + dims))
        PhiL, A, PhiR, x = _operands(*dims, seed=10 + i)
        y_ws = _with_workspace(lib, _lib, PhiL, A, PhiR, x, dims, stream)
        y = torch.full((nvec, l1, n, l2), float("nan"), dtype=torch.float64, device="cuda")
        code = lib.ttamen_local_matvec_noalloc(_p(PhiL), _p(A), _p(PhiR), _p(x), _p(y), *dims,
                                               ctypes.c_void_p(stream))
        _lib.check("ttamen_local_matvec_noalloc", code)
        torch.cuda.synchronize()
        ref = torch.einsum("lsr,smnS,LSR,vrnR->vlmL", PhiL, A, PhiR, x)
        err = float((y - ref).norm() / ref.norm())
        _say("  bitwise equal to the workspace call: %s, rel. error vs torch %.2e" % (torch.equal(y, y_ws), err))
        assert torch.equal(y, y_ws)
        assert err <= 1e-12
    # above the limit: refused before any device work (y keeps its sentinel)
    dims = (140, 4, 140, 64, 4, 140, 140, 1)
    PhiL, A, PhiR, x = _operands(*dims, seed=20)
    y = torch.full((1, 140, 64, 140), 7.0, dtype=torch.float64, device="cuda")
    code = lib.ttamen_local_matvec_noalloc(_p(PhiL), _p(A), _p(PhiR), _p(x), _p(y), *dims,
                                           ctypes.c_void_p(stream))
    torch.cuda.synchronize()
    msg = lib.ttamen_last_error().decode()
    _say("  oversize call: status %d (%s)" % (code, msg))
    assert code == 2 and "scratch" in msg
    assert bool((y == 7.0).all())
