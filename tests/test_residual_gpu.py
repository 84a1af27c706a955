"""The native global residual (solvers.residual_norms, libttamen ttamen_tt_residual_norm) against
the oracle's TT residual (a QR-based norm of the assembled residual TT), on the reference's own
solutions, on an exact cancellation (b = (1 + delta) A x: the relative residual is delta / (1 + delta)
exactly), and at d=10, n=64 where the residual TT has rank rA rx + rb > 140."""
import sys
import time

import numpy as np
import pytest
import torch

import amen_oracle as O
from conftest import GOLDEN_CASES, gpu_available, load_golden

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _tt(cores):
    import torchtt_amd as tt
    return tt.TT([torch.as_tensor(np.ascontiguousarray(c)).cuda() for c in cores])


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_residual_of_reference_solutions(name):
    import torchtt_amd as tt
    c = load_golden(name)
    r, nb = tt.solvers.residual_norms(_tt(c["A"]), _tt(c["x"]), _tt(c["b"]))
    want = O.tt_residual(c["A"], c["x"], c["b"])
    _say("%s: residual %.6e (oracle %.6e), ||b|| %.6e" % (name, r / nb, want, nb))
    assert abs(nb - O.tt_norm(c["b"])) <= 1e-13 * nb
    # the two routes differ by rounding of size u ||A x|| / ||b|| in the relative residual
    assert abs(r / nb - want) <= 1e-12 + 1e-6 * want


@pytest.mark.parametrize("delta", [1e-6, 1e-9, 1e-11])
def test_residual_exact_cancellation(delta):
    import torchtt_amd as tt
    rng = np.random.default_rng(3)
    d, n = 5, 7
    rA, rx = [1, 3, 3, 3, 3, 1], [1, 4, 5, 5, 4, 1]
    A = [rng.standard_normal((rA[k], n, n, rA[k + 1])) for k in range(d)]
    x = [rng.standard_normal((rx[k], n, rx[k + 1])) for k in range(d)]
    b = O.tt_matvec(A, x)
    b[0] = b[0] * (1.0 + delta)
    r, nb = tt.solvers.residual_norms(_tt(A), _tt(x), _tt(b))
    exact = delta / (1.0 + delta)
    _say("delta %.0e: residual %.9e exact %.9e" % (delta, r / nb, exact))
    # backward stable: error of size u ||b|| in ||A x - b||
    assert abs(r / nb - exact) <= 1e-15 + 1e-6 * exact


def test_residual_wide_residual_tt():
    """d=10, n=64 Laplacian + potential solved on the GPU: the residual TT has rank 4 rx + 4 (212 at
    the solution's rank 52), above the 140 columns of the LDS-resident QR."""
    import torchtt_amd as tt
    from torchtt_amd import workloads as W
    A, b = W.laplace_potential(10, 64, seed=0)
    Ag, bg = tt.TT(A).cuda(), tt.TT(b).cuda()
    torch.manual_seed(0)
    x = tt.solvers.amen_solve(Ag, bg, eps=1e-8, rmax=80)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r, nb = tt.solvers.residual_norms(Ag, x, bg)
    dt = time.perf_counter() - t0
    want = O.tt_residual([a.numpy() for a in A], [c.cpu().numpy() for c in x.cores], [c.numpy() for c in b])
    _say("d=10 n=64: ranks %s (residual TT rank up to %d), residual %.6e (oracle %.6e), %.1f ms" %
         (x.R, 4 * max(x.R) + 4, r / nb, want, dt * 1e3))
    assert 4 * max(x.R) + 4 > 140
    assert abs(r / nb - want) <= 1e-12 + 1e-6 * want
    assert r / nb < 1e-8
