"""Local preconditioners 'c' and 'r' (reference cpp/matvecs.h:153-194) on the GPU: the
application against the oracle's restatement, and AMEn solves that use them in the iterative
branch (the reference's own tests, tests/test_solvers.py:23-38, only reach the direct solver at
their sizes, so max_full=1 forces GMRES here; their random test operator does not converge
with GMRES at all, with or without a preconditioner, so the solves use Laplacian + potential
systems, checked against the oracle)."""
import ctypes

import numpy as np
import pytest
import torch

from conftest import gpu_available
import amen_oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _t(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("mode,shape", [("c", (4, 3, 7, 2, 5)), ("r", (3, 2, 6, 3, 5)), ("c", (9, 4, 64, 4, 11))])
def test_prec_apply_matches_oracle(mode, shape):
    import torchtt_amd as tt
    lib = tt._lib.load()
    r1, s, n, S, r2 = shape
    rng = np.random.default_rng(sum(shape))
    A = rng.standard_normal((s, n, n, S))
    for si in range(s):
        for Si in range(S):
            A[si, :, :, Si] += 3.0 * np.eye(n)
    PL = rng.standard_normal((r1, s, r1)) + 2.0
    PR = rng.standard_normal((r2, S, r2)) + 2.0
    x = rng.standard_normal((r1, n, r2))
    ref = O.prec_apply(mode, O.prec_blocks(mode, PL, A, PR), x)
    dPL, dA, dPR, dx = _t(PL), _t(A), _t(PR), _t(x)
    y = torch.empty_like(dx)
    ws = torch.empty(64 << 20, dtype=torch.uint8, device="cuda")
    code = lib.ttamen_local_prec(1 if mode == "c" else 2, ctypes.c_void_p(dPL.data_ptr()),
                                 ctypes.c_void_p(dA.data_ptr()), ctypes.c_void_p(dPR.data_ptr()),
                                 ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                 r1, s, n, S, r2, ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    tt._lib.check("ttamen_local_prec", code)
    got = y.cpu().numpy()
    assert np.linalg.norm(got - ref) <= 1e-10 * np.linalg.norm(ref)


@pytest.mark.parametrize("prec,d,n", [("c", 4, 12), ("r", 5, 6)])
def test_amen_solve_with_preconditioner(prec, d, n):
    """Same seeded inputs and residual-proxy draws as the oracle: same final ranks, solutions
    within 1e-8, residual within the reference test's bound."""
    import torchtt_amd as tt
    from torchtt_amd import workloads as W
    A, b = W.laplace_potential(d, n, seed=1, rhs_rank=2)
    torch.manual_seed(0)
    z = tt.solvers._draw_z([c.shape[1] for c in b], 4, 0, "cpu")
    xo = O.amen_solve([c.numpy() for c in A], [c.numpy() for c in b], z0=[c.numpy().copy() for c in z],
                      eps=1e-10, rmax=40, max_full=1, preconditioner=prec)
    Ag, bg = tt.TT([c.cuda() for c in A]), tt.TT([c.cuda() for c in b])
    x = tt.solvers.amen_solve(Ag, bg, eps=1e-10, rmax=40, max_full=1, preconditioner=prec, _z0=z)
    res = tt.solvers.residual(Ag, x, bg)
    assert res < 5e-8, res
    assert [c.shape[2] for c in xo] == list(x.R[1:]), (x.R, [c.shape for c in xo])
    fo = O.tt_full(xo)
    fg = O.tt_full([c.cpu().numpy() for c in x.cores])
    assert np.linalg.norm(fg - fo) <= 1e-8 * np.linalg.norm(fo)


def _system(seed):
    import torchtt_amd as tt
    from torchtt_amd import workloads as W
    A, b = W.laplace_potential(3, 6, seed=seed, rhs_rank=1)
    return tt.TT([c.cuda() for c in A]), tt.TT([c.cuda() for c in b])


def test_preconditioner_errors():
    import torchtt_amd as tt
    A, b = _system(6)
    with pytest.raises(Exception):
        tt.solvers.amen_solve(A, b, preconditioner="x")
    with pytest.raises(Exception):
        tt.solvers.amen_solve(A, b, preconditioner="c", precision="fp32")


@pytest.mark.parametrize("name,mode", [("prec_c_d4_n12", "c"), ("prec_r_d5_n6", "r")])
def test_preconditioned_solve_against_reference_fixture(name, mode):
    """GPU solve with the reference's preconditioner on the reference's own fixture (written by its
    C++ solver with the same draws, every local problem iterative): same final ranks, solution
    within 1e-8, residual within the reference test's bound (tests/test_solvers.py:23-38)."""
    import torchtt_amd as tt
    from conftest import load_golden
    c = load_golden(name)
    A = tt.TT([torch.from_numpy(a) for a in c["A"]])
    b = tt.TT([torch.from_numpy(v) for v in c["b"]])
    z = [torch.from_numpy(v.copy()) for v in c["z"]]
    x = tt.solvers.amen_solve(A.cuda(), b.cuda(), nswp=int(c["nswp"]), eps=float(c["eps"]), rmax=int(c["rmax"]),
                              max_full=1, preconditioner=mode, _z0=z)
    assert list(x.R) == c["ranks"].tolist(), (x.R, c["ranks"].tolist())
    fo = O.tt_full(c["x"])
    fg = O.tt_full([v.cpu().numpy() for v in x.cores])
    diff = np.linalg.norm(fg - fo) / np.linalg.norm(fo)
    res = tt.solvers.residual(A.cuda(), x, b.cuda())
    print("%s: rel diff vs reference %.2e residual %.2e" % (name, diff, res), flush=True)
    assert diff <= 1e-8
    assert res < 5e-8
