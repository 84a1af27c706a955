"""DMRG matvec on the GPU (torchtt_amd TT.fast_matvec / _dmrg.dmrg_matvec) against the exact
products of the reference's own tests (tests/test_algebra_2.py:39-82, in FP64) and against the
oracle's restatement of cpp/dmrg_mv.h on the same initial guess and kick stream.

Two implementations of the sweep agree on the product to the truncation tolerance and on the
final ranks (the last sweep's truncation rank of A x plus the kick); their cores are different
bases of the same TT (QR and SVD factors are unique only up to signs and, for rank-deficient
cores, rotations of the null directions), and their sweep counts may differ by one."""
import sys

import numpy as np
import pytest
import torch

import amen_oracle as O
from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel(x, y):
    return float((x - y).norm() / y.norm())


@pytest.mark.parametrize("extra", [0, 2], ids=["square", "non_square"])
def test_fast_matvec_reference(extra):
    """(A + A) fast_matvec (4 x) against 8 A x, n = 32, d = 8 (tests/test_algebra_2.py:39-82)."""
    import torchtt_amd as tt
    torch.manual_seed(0)
    n = 32
    A = tt.random([(n + extra, n)] * 8, [1] + 7 * [3] + [1], device="cuda")
    Am = A + A
    x = tt.random([n] * 8, [1] + 7 * [5] + [1], device="cuda")
    xm = x + x
    xm = xm + xm
    y = 8 * (A @ x).round(1e-12)
    _say("fast_matvec %s: start" % ("square" if extra == 0 else "non_square"))
    yf = Am.fast_matvec(xm)
    e = _rel(yf, y)
    _say("  rel err %.3e ranks %s stats %s" % (e, yf.R, tt._dmrg.last_stats))
    assert e < 1e-12


def _np_cores(t):
    return [c.detach().cpu().numpy() for c in t.cores]


def _case(name):
    rng = np.random.default_rng({"small": 1, "factored": 2, "rect": 3}[name])
    if name == "small":
        d, M, N, rA, rx = 4, 3, 3, 2, 3
    elif name == "factored":
        d, M, N, rA, rx = 5, 8, 8, 3, 4
    else:
        d, M, N, rA, rx = 5, 6, 4, 2, 3
    A = [rng.standard_normal((1 if k == 0 else rA, M, N, 1 if k == d - 1 else rA)) for k in range(d)]
    x = [rng.standard_normal((1 if k == 0 else rx, N, 1 if k == d - 1 else rx)) for k in range(d)]
    y0 = [rng.standard_normal((1 if k == 0 else 2, M, 1 if k == d - 1 else 2)) for k in range(d)]
    return A, x, y0


@pytest.mark.parametrize("name", ["small", "factored", "rect"])
def test_dmrg_matches_oracle(name):
    """Same initial guess and kick stream: equal final ranks, products within 1e-11 (the
    'factored' case takes the SVD through the factors of W at its interior supercores)."""
    import torchtt_amd as tt
    from torchtt_amd._dmrg import dmrg_matvec
    A, x, y0 = _case(name)
    tr = {}
    y_or = O.dmrg_mv(A, x, y0, nswp=20, eps=1e-12, kickrank=4, seed=77, trace=tr)
    to = lambda cs: tt.TT([torch.tensor(c, device="cuda") for c in cs])
    _say("dmrg %s: start" % name)
    y = dmrg_matvec(to(A), to(x), y0=to(y0), nswp=20, eps=1e-12, kickrank=4, _seed=77)
    st = tt._dmrg.last_stats
    ref = to(y_or)
    e = _rel(y, ref)
    _say("  ranks gpu %s oracle %s, sweeps gpu %d oracle %d, diff %.3e" % (
        list(y.R), [1] + [c.shape[2] for c in y_or], st["sweeps"], tr["sweeps"], e))
    assert list(y.R) == [1] + [c.shape[2] for c in y_or]
    assert abs(st["sweeps"] - tr["sweeps"]) <= 1
    assert e < 1e-11
    exact = O.tt_matvec(A, x)
    ee = O.tt_norm(O.tt_add(_np_cores(y), exact, 1.0, -1.0)) / O.tt_norm(exact)
    assert ee < 1e-12


def test_dmrg_identity_zero_and_caps():
    import torchtt_amd as tt
    torch.manual_seed(3)
    N = [4, 5, 3, 4]
    x = tt.random(N, [1, 3, 4, 2, 1], device="cuda")
    y = tt.eye(N, device="cuda").fast_matvec(x)
    assert _rel(y, x) < 1e-13
    z = tt.zeros(N, device="cuda")
    y = tt.eye(N, device="cuda").fast_matvec(z, nswp=4)
    assert list(y.N) == N and float(y.norm()) == 0.0
    from torchtt_amd._dmrg import dmrg_matvec
    A = tt.random([(n, n) for n in N], [1, 3, 3, 3, 1], device="cuda")
    y = dmrg_matvec(A, x, rmax=5, kickrank=2, nswp=6)
    assert max(y.R) <= 5 + 2


def test_dmrg_errors():
    import torchtt_amd as tt
    from torchtt_amd._dmrg import dmrg_matvec
    from torchtt_amd.errors import InvalidArguments, IncompatibleTypes, ShapeMismatch
    A = tt.random([(4, 4)] * 3, [1, 2, 2, 1], device="cuda")
    x = tt.random([4] * 3, [1, 2, 2, 1], device="cuda")
    with pytest.raises(IncompatibleTypes):
        x.fast_matvec(x)
    with pytest.raises(ShapeMismatch):
        A.fast_matvec(tt.random([4, 4, 5], [1, 2, 2, 1], device="cuda"))
    with pytest.raises(InvalidArguments):
        tt.random([(4, 4)], [1, 1], device="cuda").fast_matvec(tt.random([4], [1, 1], device="cuda"))
    with pytest.raises(InvalidArguments):
        dmrg_matvec(A, tt.TT([c.to(torch.complex128) for c in x.cores]))
    # an SVD beyond the LDS-resident size: at supercore (1, 2), W is 256 x 160 of inner rank 150
    xb = tt.random([16] * 4, [1, 16, 150, 16, 1], device="cuda")
    Ab = tt.eye([16] * 4, device="cuda")
    y0 = tt.random([16] * 4, [1, 10, 10, 10, 1], device="cuda")
    with pytest.raises(Exception, match="not supported"):
        dmrg_matvec(Ab, xb, y0=y0)
