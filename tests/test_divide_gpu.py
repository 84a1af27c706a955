"""Elementwise division by AMEn on the GPU (torchtt_amd.solvers.amen_divide, TT.__truediv__,
elementwise_divide) against the dense quotient -- the reference's own tests
(tests/test_algebra_2.py:85-120) -- and against the oracle's restatement of
torchtt/_division.py on the same seeded residual-proxy draws."""
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


def _err_rel(t, ref):
    return float(torch.linalg.norm(t - ref) / torch.linalg.norm(ref))


def _reference_operands(dev):
    """x = x0 + x1 + x2 + x3 + x1 x2 + (1 - x3) x2 + 1 on a 7x8x9x10 grid of [0,1]^4 (rounded), y = 1."""
    import torchtt_amd as tt
    N = [7, 8, 9, 10]
    xs = tt.meshgrid([torch.linspace(0, 1, n, dtype=torch.float64, device=dev) for n in N])
    x = xs[0] + xs[1] + xs[2] + xs[3] + xs[1] * xs[2] + (1 - xs[3]) * xs[2] + 1
    x = x.round(1e-14)
    y = tt.ones(x.N, dtype=torch.float64, device=dev)
    return x, y


def test_division_operators():
    """y / x, 1 / x and tensor(1.0) / x (reference tests/test_algebra_2.py:85-102)."""
    _say("divide: operators")
    x, y = _reference_operands("cuda")
    a = y / x
    b = 1 / x
    c = torch.tensor(1.0) / x
    ref = y.full() / x.full()
    for name, t in (("y/x", a), ("1/x", b), ("1.0/x", c)):
        e = _err_rel(t.full(), ref)
        _say("  %s rel err %.3e ranks %s" % (name, e, t.R))
        assert e < 1e-11


def test_division_preconditioned():
    """elementwise_divide with the local preconditioner 'c' (tests/test_algebra_2.py:104-120)."""
    import torchtt_amd as tt
    _say("divide: preconditioned")
    x, y = _reference_operands("cuda")
    a = tt.elementwise_divide(y, x, preconditioner='c')
    e = _err_rel(a.full(), y.full() / x.full())
    _say("  rel err %.3e ranks %s" % (e, a.R))
    assert e < 1e-11


@pytest.mark.parametrize("case", ["reference_grid", "random_positive"])
def test_divide_matches_oracle(case):
    """Same operands and residual-proxy draws as the oracle: same quotient within 1e-9 of each
    other and of the dense quotient to the accuracy asked for."""
    import torchtt_amd as tt
    _say("divide vs oracle", case)
    if case == "reference_grid":
        x, y = _reference_operands("cpu")
        a_np, b_np = [c.numpy() for c in x.cores], [c.numpy() for c in y.cores]
        eps, rmax = 1e-12, 500
    else:
        rng = np.random.default_rng(4)
        N = [6, 9, 7, 8, 5]
        ra = [1, 2, 3, 3, 2, 1]
        a_np = [0.2 * rng.standard_normal((ra[k], N[k], ra[k + 1])) for k in range(5)]
        a_np[0][0, :, 0] += 3.0  # keep a away from zero: a = 3 + small rank-3 part
        a_np = O.tt_add([np.full((1, n, 1), 1.0) * (3.0 if k == 0 else 1.0) for k, n in enumerate(N)],
                        [c.copy() for c in a_np])
        rb = [1, 2, 2, 2, 2, 1]
        b_np = [rng.standard_normal((rb[k], N[k], rb[k + 1])) for k in range(5)]
        eps, rmax = 1e-10, 40
    torch.manual_seed(3)
    N = [c.shape[1] for c in a_np]
    d = len(N)
    rz = [1] + [4] * (d - 1) + [1]
    z0 = [torch.randn((rz[k], N[k], rz[k + 1]), dtype=torch.float64) for k in range(d)]
    tr = {}
    q_or = O.amen_divide(a_np, b_np, z0=[z.numpy().copy() for z in z0], nswp=50, eps=eps, rmax=rmax, trace=tr)
    a = tt.TT([torch.tensor(c).cuda() for c in a_np])
    b = tt.TT([torch.tensor(c).cuda() for c in b_np])
    torch.manual_seed(3)
    q = tt.TT(tt.solvers.amen_divide(a, b, nswp=50, eps=eps, rmax=rmax, verbose=False))
    full_q = q.full().cpu().numpy()
    full_o = O.tt_full(q_or)
    want = O.tt_full(b_np) / O.tt_full(a_np)
    d_or = np.linalg.norm(full_q - full_o) / np.linalg.norm(full_o)
    e = np.linalg.norm(full_q - want) / np.linalg.norm(want)
    _say("  ranks gpu %s oracle %s, diff vs oracle %.3e, err vs dense %.3e, stats %s"
         % (q.R, [1] + [c.shape[2] for c in q_or], d_or, e, tt.solvers.last_stats))
    assert d_or < 1e-9
    assert e < 50 * eps
    # Ranks within one of the oracle's at each interface: at eps = 1e-12 the truncation compares local
    # residuals at the rounding level, and the oracle's own middle rank on the reference grid moves
    # between 29 and 30 when the operand is perturbed by 2e-16 relative (a summation order does as much).
    r_or = [1] + [c.shape[2] for c in q_or]
    assert len(q.R) == len(r_or) and all(abs(u - v) <= 1 for u, v in zip(q.R, r_or)), (q.R, r_or)
    # the stop test max_res < eps compares local residuals that GMRES drives to ~eps/(2 sqrt(d)):
    # one sweep more or less is within the rounding of two implementations (same ranks, same quotient)
    assert abs(tt.solvers.last_stats["sweeps"] - tr["sweeps"]) <= 1


def test_divide_errors():
    import torchtt_amd as tt
    from torchtt_amd.errors import ShapeMismatch, InvalidArguments
    x = tt.ones([3, 4], dtype=torch.float64, device="cuda")
    y = tt.ones([3, 5], dtype=torch.float64, device="cuda")
    with pytest.raises(ShapeMismatch):
        tt.solvers.amen_divide(x, y)
    with pytest.raises(InvalidArguments):
        tt.solvers.amen_divide(x, x, preconditioner='r')
    with pytest.raises(InvalidArguments):
        tt.solvers.amen_divide(x, x, rmax=[1, 2])
