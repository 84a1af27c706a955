"""AMEn products on the GPU (torchtt_amd.amen_mv / amen_mm) against the exact products -- the
reference's own tests (tests/test_algebra_2.py:122-245) -- and against the oracle's restatement
of torchtt/_amen.py on the same residual-proxy draws."""
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


def _five(t):
    return t + t + t + t + t


def test_amen_mv_product():
    """5A times 5x against 25 A x (tests/test_algebra_2.py:122-138)."""
    import torchtt_amd as tt
    torch.manual_seed(0)
    A = tt.randn([(3, 4), (5, 6), (7, 8), (2, 3)], [1, 2, 2, 3, 1], device="cuda")
    x = tt.randn([4, 6, 8, 3], [1, 4, 3, 3, 1], device="cuda")
    Cr = 25 * (A @ x)
    C = tt.amen_mv(_five(A), _five(x))
    e = _rel(C, Cr)
    _say("amen_mv: rel err %.3e ranks %s stats %s" % (e, C.R, tt._amen.last_stats))
    assert e < 1e-11


def test_amen_mv_special_operands():
    """Zero operator, zero vector, identity (tests/test_algebra_2.py:141-185)."""
    import torchtt_amd as tt
    N = [3, 4, 2]
    torch.manual_seed(1)
    x = tt.randn(N, [1, 2, 2, 1], device="cuda")
    C = tt.amen_mv(tt.zeros([(n, n) for n in N], device="cuda"), x, nswp=4, eps=1e-12, kickrank=2)
    assert list(C.N) == N and float(C.norm()) < 1e-12
    C = tt.amen_mv(tt.eye(N, device="cuda"), tt.zeros(N, device="cuda"), nswp=4, eps=1e-12, kickrank=2)
    assert list(C.N) == N and float(C.norm()) < 1e-12
    C = tt.amen_mv(tt.eye(N, device="cuda"), x, nswp=8, eps=1e-12, kickrank=2)
    e = _rel(C, x)
    _say("amen_mv identity: rel err %.3e" % e)
    assert e < 1e-11


def test_amen_mv_rank_above_140():
    """A product whose rank outgrows the default rank capacity (140): the sweep stops with status 4
    and runs again at a larger capacity, its QR and SVD above 140 columns in device memory. Identity
    times a random TT of rank 180 at the middle interface (16^4: at most 256 there), warm-started from
    that TT (the plan holds 180 there; the enrichment asks for 184), reproduces it."""
    import torchtt_amd as tt
    g = torch.Generator().manual_seed(11)
    N, r = [16] * 4, [1, 16, 180, 16, 1]
    x = tt.TT([torch.randn(r[k], N[k], r[k + 1], dtype=torch.float64, generator=g).cuda() for k in range(4)])
    C = tt.amen_mv(tt.eye(N, device="cuda"), x, x0=x, nswp=2, eps=1e-12)
    e = _rel(C, x)
    _say("amen_mv rank 180: rel err %.3e ranks %s" % (e, C.R))
    assert max(C.R) > 140
    assert e < 1e-10


def test_amen_mm_product():
    """5A times 5B against 25 A B (tests/test_algebra_2.py:188-204)."""
    import torchtt_amd as tt
    torch.manual_seed(2)
    A = tt.randn([(3, 4), (5, 6), (7, 8), (2, 3)], [1, 2, 2, 3, 1], device="cuda")
    B = tt.randn([(4, 2), (6, 4), (8, 5), (3, 7)], [1, 4, 3, 3, 1], device="cuda")
    Cr = 25 * (A @ B)
    C = tt.amen_mm(_five(A), _five(B))
    e = _rel(C, Cr)
    _say("amen_mm: rel err %.3e ranks %s" % (e, C.R))
    assert list(C.M) == [3, 5, 7, 2] and list(C.N) == [2, 4, 5, 7]
    assert e < 1e-11


def test_amen_mm_special_operands():
    """Zero left/right factors, identity left factor (tests/test_algebra_2.py:207-245)."""
    import torchtt_amd as tt
    torch.manual_seed(3)
    B = tt.randn([(2, 5), (3, 6)], [1, 2, 1], device="cuda")
    C = tt.amen_mm(tt.zeros([(3, 2), (4, 3)], device="cuda"), B, nswp=4, eps=1e-12, kickrank=2)
    assert list(C.M) == [3, 4] and list(C.N) == [5, 6] and float(C.norm()) < 1e-12
    C = tt.amen_mm(tt.eye([3, 4], device="cuda"), tt.zeros([(3, 2), (4, 5)], device="cuda"), nswp=4, eps=1e-12,
                   kickrank=2)
    assert list(C.M) == [3, 4] and list(C.N) == [2, 5] and float(C.norm()) < 1e-12
    B2 = tt.randn([(3, 2), (4, 5)], [1, 2, 1], device="cuda")
    C = tt.amen_mm(tt.eye([3, 4], device="cuda"), B2, nswp=8, eps=1e-12, kickrank=2)
    assert _rel(C, B2) < 1e-11


@pytest.mark.parametrize("case", ["product", "rectangular_low_rank"])
def test_amen_mv_matches_oracle(case):
    """Same operands and residual-proxy draws as the oracle: same ranks and sweeps, products
    within 1e-10 of each other."""
    import torchtt_amd as tt
    rng = np.random.default_rng(7)
    if case == "product":
        shape, RA, Rx = [(3, 4), (5, 6), (7, 8), (2, 3)], [1, 2, 2, 3, 1], [1, 4, 3, 3, 1]
    else:
        shape, RA, Rx = [(6, 4), (5, 9), (8, 7), (4, 6), (7, 5)], [1, 3, 2, 3, 2, 1], [1, 2, 3, 3, 2, 1]
    d = len(shape)
    A = [rng.standard_normal((RA[k], shape[k][0], shape[k][1], RA[k + 1])) / 3 for k in range(d)]
    x = [rng.standard_normal((Rx[k], shape[k][1], Rx[k + 1])) / 3 for k in range(d)]
    torch.manual_seed(11)
    rz = [1] + [4] * (d - 1) + [1]
    z0 = [torch.randn((rz[k], shape[k][0], rz[k + 1]), dtype=torch.float64) for k in range(d)]
    tr = {}
    y_or = O.amen_mv(A, x, z0=[z.numpy().copy() for z in z0], eps=1e-10, trace=tr)
    torch.manual_seed(11)
    y = tt.amen_mv(tt.TT([torch.tensor(c) for c in A]), tt.TT([torch.tensor(c) for c in x]), eps=1e-10)
    fy, fo = y.full().numpy(), O.tt_full(y_or)
    exact = O.tt_full(O.tt_matvec(A, x))
    diff = np.linalg.norm(fy - fo) / np.linalg.norm(fo)
    err = np.linalg.norm(fy - exact) / np.linalg.norm(exact)
    st = tt._amen.last_stats
    _say("amen_mv vs oracle %s: ranks gpu %s oracle %s, sweeps %d/%d, diff %.3e, err vs exact %.3e"
         % (case, y.R, [1] + [c.shape[2] for c in y_or], st["sweeps"], tr["sweeps"], diff, err))
    assert list(y.R) == [1] + [c.shape[2] for c in y_or]
    assert st["sweeps"] == tr["sweeps"]
    assert diff < 1e-10 and err < 1e-9


def test_amen_products_errors():
    import torchtt_amd as tt
    from torchtt_amd.errors import IncompatibleTypes, ShapeMismatch, InvalidArguments
    A = tt.eye([3, 4], device="cuda")
    with pytest.raises(IncompatibleTypes):
        tt.amen_mv(A, A)
    with pytest.raises(ShapeMismatch):
        tt.amen_mv(A, tt.ones([3, 5], device="cuda"))
    with pytest.raises(InvalidArguments):
        tt.amen_mm(A, tt.ones([3, 4], device="cuda"))
    with pytest.raises(InvalidArguments):
        tt.amen_mv(A, tt.ones([3, 4], device="cuda"), rmax=[1, 2])
