"""End-to-end parity of the GPU AMEn solver (torchtt_amd.solvers.amen_solve).

Bars: the solution agrees with the reference solver's solution on the same seeded inputs
to a relative Frobenius difference <= 1e-8 and has the same final TT ranks; the relative
TT residual ||Ax - b|| / ||b|| is checked for every result.
"""
import sys

import numpy as np
import pytest
import torch

import amen_oracle as O
import laplace_exact as LE
from conftest import GOLDEN_CASES, GOLDEN_SWEEPS, gpu_available, load_golden, load_ref

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _tt(cores, dev="cuda"):
    import torchtt_amd as tt
    return tt.TT([torch.as_tensor(np.ascontiguousarray(c)).to(dev) for c in cores])


def _rel_tt(x, y):
    """||x - y|| / ||y|| in TT arithmetic (QR-based norm, no cancellation)."""
    return float((x - y).norm() / y.norm())


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_amen_matches_reference_fixture(name):
    import torchtt_amd as tt
    c = load_golden(name)
    _say("amen golden", name)
    A, b = _tt(c["A"]), _tt(c["b"])
    x0 = _tt(c["x0"]) if c["x0"] is not None else None
    torch.manual_seed(c["seed"])
    # the residual-proxy cores are drawn from the CPU generator, as the reference did
    A_cpu = _tt(c["A"], "cpu")
    z_expected = [torch.from_numpy(z) for z in c["z"]]
    torch.manual_seed(c["seed"])
    x = tt.solvers.amen_solve(A_cpu, _tt(c["b"], "cpu"), x0=_tt(c["x0"], "cpu") if x0 is not None else None,
                              nswp=c["nswp"], eps=c["eps"], rmax=c["rmax"], kickrank=c["kickrank"], kick2=c["kick2"])
    # check the draws we fed the solver are the fixture's
    torch.manual_seed(c["seed"])
    d = c["d"]
    rz = [1] + [c["kickrank"] + c["kick2"]] * (d - 1) + [1]
    for k in range(d):
        zk = torch.randn((rz[k], A_cpu.N[k], rz[k + 1]), dtype=torch.float64)
        assert torch.equal(zk, z_expected[k])
    xr = _tt(c["x"], "cpu")
    assert x.R == c["ranks"].tolist(), (x.R, c["ranks"].tolist())
    diff = _rel_tt(x, xr)
    res = tt.solvers.residual(A_cpu, x, _tt(c["b"], "cpu"))
    _say("  rel diff vs reference %.3e, residual %.3e, stats %s" % (diff, res, tt.solvers.last_stats))
    assert diff <= 1e-8
    assert res <= 50 * c["eps"]
    assert tt.solvers.last_stats["sweeps"] == GOLDEN_SWEEPS[name], tt.solvers.last_stats


def test_laplace_closed_form():
    """d=6, n=16, h=1/17 Laplacian, b = ones, rank cap 20: error against the exact solution."""
    import torchtt_amd as tt
    from torchtt_amd import workloads as W
    _say("amen laplace closed form")
    A, b = W.laplace_ones(6, 16)
    A, b = tt.TT([c.cuda() for c in A]), tt.TT([c.cuda() for c in b])
    torch.manual_seed(0)
    x = tt.solvers.amen_solve(A, b, nswp=10, eps=1e-10, rmax=20)
    ex = LE.laplace_solve_full(6, 16, 1.0 / 17)
    err = np.linalg.norm(x.full().cpu().numpy() - ex) / np.linalg.norm(ex)
    res = tt.solvers.residual(A, x, b)
    _say("  error vs closed form %.3e residual %.3e" % (err, res))
    assert err < 1e-10
    assert res < 1e-9


def test_laplace_potential_d10_against_reference():
    """d=10, n=64, Laplacian + potential (rank 4), random rank-4 rhs, rank cap 80."""
    import torchtt_amd as tt
    from torchtt_amd import workloads as W
    ref = load_ref()
    _say("amen d10 n64 vs", "reference build" if ref is not None else "oracle")
    A, b = W.laplace_potential(10, 64, seed=0)
    eps = 1e-10
    torch.manual_seed(0)
    x = tt.solvers.amen_solve(tt.TT(A), tt.TT(b), eps=eps, rmax=80)
    stats = dict(tt.solvers.last_stats)
    torch.manual_seed(0)
    if ref is not None:
        torch.set_num_threads(min(16, torch.get_num_threads()))
        xr = ref.amen_solve(A, b, [], [64] * 10, tt.TT(A).R, tt.TT(b).R, [1] * 11, 22, eps, 80, 500, 4, 0, 40, 2,
                            False, 0)
        xr = tt.TT(list(xr))
    else:
        rz = [1] + [4] * 9 + [1]
        z0 = [torch.randn((rz[k], 64, rz[k + 1]), dtype=torch.float64).numpy() for k in range(10)]
        xr = tt.TT([torch.from_numpy(c) for c in O.amen_solve([a.numpy() for a in A], [c.numpy() for c in b],
                                                              z0=z0, eps=eps, rmax=80)])
    diff = _rel_tt(x, xr)
    res = tt.solvers.residual(tt.TT(A), x, tt.TT(b))
    _say("  ranks gpu %s ref %s" % (x.R, xr.R))
    _say("  rel diff %.3e residual %.3e stats %s" % (diff, res, stats))
    assert res < 1e-8
    assert diff < 1e-8


@pytest.mark.parametrize("name", ["potential_d4_n16", "laplace_d6_n16_eps1e-10"])
def test_mixed_precision_against_reference_fixture(name):
    """FP32 local GMRES inside FP64 refinement reaches the same solution as the FP64 reference."""
    import torchtt_amd as tt
    c = load_golden(name)
    _say("amen fp32+refinement golden", name)
    torch.manual_seed(c["seed"])
    x = tt.solvers.amen_solve(_tt(c["A"], "cpu"), _tt(c["b"], "cpu"), nswp=c["nswp"], eps=c["eps"],
                              rmax=c["rmax"], kickrank=c["kickrank"], precision="fp32")
    st = dict(tt.solvers.last_stats)
    xr = _tt(c["x"], "cpu")
    diff = _rel_tt(x, xr)
    res = tt.solvers.residual(_tt(c["A"], "cpu"), x, _tt(c["b"], "cpu"))
    _say("  ranks %s (ref %s) rel diff vs FP64 reference %.3e residual %.3e refinements %d fp32 matvecs %d"
         % (x.R, c["ranks"].tolist(), diff, res, st["refinements"], st["matvecs_f32"]))
    assert st["matvecs_f32"] > 0
    assert res <= 50 * c["eps"]
    assert diff <= 1e-8


@pytest.mark.parametrize("name", ["potential_d4_n16", "laplace_d6_n16_eps1e-10"])
def test_bf16_against_reference_fixture(name):
    """BF16 products (FP32 accumulation) inside FP64 refinement reach the FP64 reference's solution:
    the refinement residual is FP64, so only the number of refinement steps depends on BF16."""
    import torchtt_amd as tt
    c = load_golden(name)
    _say("amen bf16+refinement golden", name)
    torch.manual_seed(c["seed"])
    x = tt.solvers.amen_solve(_tt(c["A"], "cpu"), _tt(c["b"], "cpu"), nswp=c["nswp"], eps=c["eps"],
                              rmax=c["rmax"], kickrank=c["kickrank"], precision="bf16")
    st = dict(tt.solvers.last_stats)
    diff = _rel_tt(x, _tt(c["x"], "cpu"))
    res = tt.solvers.residual(_tt(c["A"], "cpu"), x, _tt(c["b"], "cpu"))
    _say("  ranks %s (ref %s) rel diff vs FP64 reference %.3e residual %.3e refinements %d bf16 matvecs %d"
         % (x.R, c["ranks"].tolist(), diff, res, st["refinements"], st["matvecs_f32"]))
    assert st["matvecs_f32"] > 0 and st["refinements"] > 0
    assert res <= 50 * c["eps"]
    assert diff <= 1e-7


def test_mixed_precision_d8_n48():
    """Laplacian + potential d=8, n=48, rank cap 60: FP32 mode vs FP64 mode on the GPU."""
    import torchtt_amd as tt
    from torchtt_amd import workloads as W
    A, b = W.laplace_potential(8, 48, seed=2)
    A, b = tt.TT(A).cuda(), tt.TT(b).cuda()
    _say("amen d8 n48 fp64 vs fp32")
    torch.manual_seed(0)
    x64 = tt.solvers.amen_solve(A, b, eps=1e-10, rmax=60)
    torch.manual_seed(0)
    x32 = tt.solvers.amen_solve(A, b, eps=1e-10, rmax=60, precision="fp32")
    st = dict(tt.solvers.last_stats)
    r64, r32 = tt.solvers.residual(A, x64, b), tt.solvers.residual(A, x32, b)
    diff = _rel_tt(x32, x64)
    _say("  residual fp64 %.3e fp32 %.3e, rel diff %.3e, stats %s" % (r64, r32, diff, st))
    assert r32 < 1e-8 and r64 < 1e-8
    assert diff < 1e-7


def test_edge_cases():
    """d = 1, d = 2, no enrichment (kickrank = 0, which the reference rejects), warm start."""
    import torchtt_amd as tt
    rng = np.random.default_rng(9)
    # d = 1: a single dense system solved by the direct local solver
    A1 = [rng.standard_normal((1, 12, 12, 1)) + 12 * np.eye(12).reshape(1, 12, 12, 1)]
    b1 = [rng.standard_normal((1, 12, 1))]
    _say("amen d=1")
    x1 = tt.solvers.amen_solve(_tt(A1), _tt(b1), eps=1e-12)
    want = np.linalg.solve(A1[0][0, :, :, 0], b1[0][0, :, 0])
    assert np.linalg.norm(x1.full().cpu().numpy() - want) / np.linalg.norm(want) < 1e-12
    # d = 2
    from torchtt_amd import workloads as W
    A2, b2 = W.laplace_potential(2, 30, seed=3, rhs_rank=2)
    _say("amen d=2")
    x2 = tt.solvers.amen_solve(tt.TT(A2).cuda(), tt.TT(b2).cuda(), eps=1e-10)
    assert tt.solvers.residual(tt.TT(A2).cuda(), x2, tt.TT(b2).cuda()) < 1e-9
    # kickrank = 0: no enrichment, so ranks cannot grow past the start's (one-site ALS); the
    # reference raises on this input. From a start of ranks (10, 16, 10) the sweeps converge.
    A3, b3 = W.laplace_ones(4, 10)
    _say("amen kickrank=0")
    g = torch.Generator().manual_seed(4)
    x0 = tt.TT([torch.randn(r0, 10, r1, dtype=torch.float64, generator=g).cuda()
                for r0, r1 in zip([1, 10, 16, 10], [10, 16, 10, 1])])
    x3 = tt.solvers.amen_solve(tt.TT(A3).cuda(), tt.TT(b3).cuda(), x0=x0, eps=1e-10, kickrank=0, nswp=30)
    assert max(x3.R) <= 16
    assert tt.solvers.residual(tt.TT(A3).cuda(), x3, tt.TT(b3).cuda()) < 1e-8
    # input validation errors match the reference
    with pytest.raises(tt.errors.InvalidArguments):
        tt.solvers.amen_solve(tt.TT(A3), b3)
    with pytest.raises(tt.errors.IncompatibleTypes):
        tt.solvers.amen_solve(tt.TT(b3), tt.TT(b3))


def test_contract_of_the_return_value():
    """The reference's return contract (torchtt/solvers.py:285, amen_solve.h:157): the solution comes
    back on the device and dtype of A.cores[0] (CPU or FP32 inputs included), rank caps hold, nswp
    bounds the sweeps, mode sizes may differ per core, and b = 0 gives x = 0 with finite values."""
    import torchtt_amd as tt
    rng = np.random.default_rng(21)
    N = [3, 7, 2, 5]
    d = len(N)
    A = [rng.standard_normal((1 if k == 0 else 2, n, n, 1 if k == d - 1 else 2)) * 0.1 for k, n in enumerate(N)]
    for k, n in enumerate(N):  # a dominant diagonal keeps the system well conditioned
        A[k][0, :, :, 0] += np.eye(n) * (2.0 if k == 0 else 1.0)
    b = [rng.standard_normal((1 if k == 0 else 3, n, 1 if k == d - 1 else 3)) for k, n in enumerate(N)]
    _say("contract: CPU inputs")
    x = tt.solvers.amen_solve(_tt(A, "cpu"), _tt(b, "cpu"), eps=1e-12)
    assert x.cores[0].device.type == "cpu" and x.cores[0].dtype == torch.float64
    assert list(x.N) == N
    assert tt.solvers.residual(_tt(A), x.to(device="cuda"), _tt(b)) < 1e-10
    _say("contract: FP32 inputs")
    A32 = tt.TT([torch.tensor(c, dtype=torch.float32, device="cuda") for c in A])
    b32 = tt.TT([torch.tensor(c, dtype=torch.float32, device="cuda") for c in b])
    x32 = tt.solvers.amen_solve(A32, b32, eps=1e-10)
    assert x32.cores[0].dtype == torch.float32 and x32.cores[0].is_cuda
    _say("contract: rank cap and nswp")
    from torchtt_amd import workloads as W
    Al, bl = W.laplace_potential(6, 12, seed=2, rhs_rank=3)
    xr = tt.solvers.amen_solve(tt.TT(Al).cuda(), tt.TT(bl).cuda(), eps=1e-12, rmax=3, nswp=4)
    # truncation caps at rmax; enrichment then adds kickrank (4) in every sweep but the last one
    # (amen_solve.h:443-460), and a solve stopped by nswp never reached its last sweep
    assert max(xr.R) <= 3 + 4
    assert tt.solvers.last_stats["sweeps"] <= 4
    xs = tt.solvers.amen_solve(tt.TT(Al).cuda(), tt.TT(bl).cuda(), eps=1e-12, nswp=1)
    assert tt.solvers.last_stats["sweeps"] == 1
    assert all(torch.isfinite(c).all() for c in xs.cores)
    _say("contract: zero right-hand side")
    b0 = tt.TT([torch.zeros_like(torch.as_tensor(c)).cuda() for c in b])
    x0 = tt.solvers.amen_solve(_tt(A), b0, eps=1e-10, nswp=3)
    assert all(torch.isfinite(c).all() for c in x0.cores)
    assert float(x0.norm()) == 0.0


def test_rank_above_140_solves_at_a_raised_capacity():
    """A solve whose ranks outgrow the default rank capacity (140, the LDS-resident QR/SVD) stops with
    status 4 and is solved again at a larger capacity, where QR and SVD of more than 140 columns run
    in device memory. Laplacian + potential on 16^4, a random right-hand side of rank 200 at the middle
    interface and a warm start of rank 140 there: the enrichment asks for 144 at once. Same operands
    and residual-proxy draws as the oracle: the same ranks (within one: eps 1e-10 truncates at the
    rounding level of the local residuals) and solutions within 1e-8 of each other."""
    import torchtt_amd as tt
    from torchtt_amd import workloads as W
    g = torch.Generator().manual_seed(5)
    N, rb, r0 = [16] * 4, [1, 16, 200, 16, 1], [1, 16, 140, 16, 1]
    A, _ = W.laplace_potential(4, 16, seed=2)
    b = [torch.randn(rb[k], N[k], rb[k + 1], dtype=torch.float64, generator=g) for k in range(4)]
    x0 = [torch.randn(r0[k], N[k], r0[k + 1], dtype=torch.float64, generator=g) for k in range(4)]
    torch.manual_seed(7)
    z0 = tt.solvers._draw_z(N, 4, 0, "cpu")
    _say("rank above 140: GPU solve")
    lib = tt._lib.load()
    assert lib.ttamen_rank_capacity(0) == 140
    x = tt.solvers.amen_solve(tt.TT([c.cuda() for c in A]), tt.TT([c.cuda() for c in b]),
                              x0=tt.TT([c.cuda() for c in x0]), eps=1e-10, nswp=2, _z0=z0)
    stats = dict(tt.solvers.last_stats)
    assert lib.ttamen_rank_capacity(0) == 140  # restored
    _say("  ranks %s, capacity %d; oracle" % (x.R, stats["rank_capacity"]))
    assert max(x.R) > 140 and stats["rank_capacity"] > 140
    xo = O.amen_solve([c.numpy() for c in A], [c.numpy() for c in b], x0=[c.numpy() for c in x0],
                      z0=[z.numpy() for z in z0], nswp=2, eps=1e-10)
    ro = [1] + [c.shape[2] for c in xo]
    fx, fo = x.full().cpu().numpy(), O.tt_full(xo)
    diff = np.linalg.norm(fx - fo) / np.linalg.norm(fo)
    _say("  oracle ranks %s, rel diff %.3e" % (ro, diff))
    assert all(abs(u - v) <= 1 for u, v in zip(x.R, ro)), (x.R, ro)
    assert diff < 1e-8


def test_rank_capacity_abi():
    """ttamen_rank_capacity: per calling thread, 140 by default, raised up to 1024, above refused."""
    import torchtt_amd as tt
    lib = tt._lib.load()
    assert lib.ttamen_rank_capacity(0) == 140
    assert lib.ttamen_rank_capacity(300) == 140
    assert lib.ttamen_rank_capacity(0) == 300
    assert lib.ttamen_rank_capacity(2000) == -1
    assert lib.ttamen_rank_capacity(100) == 300   # below the LDS size: clamped to it
    assert lib.ttamen_rank_capacity(0) == 140


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fused_local_operator_solve(precision):
    """The solve with the fused local operator (ttamen_matvec_mode(1)): the reference fixture's
    ranks and solution, as with the three-GEMM form."""
    import torchtt_amd as tt
    lib = tt._lib.load()
    c = load_golden("potential_d4_n16")
    prev = lib.ttamen_matvec_mode(1)
    try:
        _say("amen golden potential_d4_n16, fused local operator,", precision)
        torch.manual_seed(c["seed"])
        x = tt.solvers.amen_solve(_tt(c["A"], "cpu"), _tt(c["b"], "cpu"), nswp=c["nswp"], eps=c["eps"],
                                  rmax=c["rmax"], kickrank=c["kickrank"], precision=precision)
    finally:
        lib.ttamen_matvec_mode(prev)
    xr = _tt(c["x"], "cpu")
    diff = _rel_tt(x, xr)
    res = tt.solvers.residual(_tt(c["A"], "cpu"), x, _tt(c["b"], "cpu"))
    _say("  ranks %s (ref %s) rel diff %.3e residual %.3e" % (x.R, c["ranks"].tolist(), diff, res))
    assert x.R == c["ranks"].tolist()
    assert diff <= 1e-8 and res <= 50 * c["eps"]
