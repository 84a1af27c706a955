"""TT rounding on the GPU (ttamen_tt_round) against the oracle's restatement of the reference's
round_this (cpp/compression.h:3-41) and against what rounding must preserve."""
import numpy as np
import pytest
import torch

from conftest import gpu_available
import amen_oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _random_cores(rng, n, ranks):
    return [rng.standard_normal((ranks[k], n[k], ranks[k + 1])) for k in range(len(n))]


def _full(cores):
    return O.tt_full([np.asarray(c.reshape(c.shape[0], -1, c.shape[-1])) for c in cores])


def _gpu_round(cores, eps, rmax):
    import torchtt_amd as tt
    x = tt.TT([torch.tensor(c, dtype=torch.float64, device="cuda") for c in cores])
    y = tt.solvers.round_tt(x, eps, rmax)
    return [c.cpu().numpy() for c in y.cores], y


@pytest.mark.parametrize("case", ["redundant", "random", "capped"])
def test_round_matches_oracle(case):
    rng = np.random.default_rng({"redundant": 1, "random": 2, "capped": 3}[case])
    n = [5, 6, 4, 7, 5]
    if case == "redundant":
        a = _random_cores(rng, n, [1, 3, 4, 4, 2, 1])
        # a + a as a TT of doubled ranks (block-diagonal cores): rounds back to a's ranks
        cores = []
        d = len(n)
        for k in range(d):
            c = a[k]
            if k == 0:
                cores.append(np.concatenate([c, c], axis=2))
            elif k == d - 1:
                cores.append(np.concatenate([c, c], axis=0))
            else:
                z = np.zeros((2 * c.shape[0], c.shape[1], 2 * c.shape[2]))
                z[:c.shape[0], :, :c.shape[2]] = c
                z[c.shape[0]:, :, c.shape[2]:] = c
                cores.append(z)
        eps, rmax = 1e-12, 1000
        exact = 2 * _full(a)
    else:
        cores = _random_cores(rng, n, [1, 4, 9, 9, 5, 1])
        eps, rmax = (1e-1, 1000) if case == "random" else (1e-14, 3)
        exact = _full(cores)
    ref = O.round_tt([c.copy() for c in cores], eps, rmax)
    got, y = _gpu_round(cores, eps, rmax)
    assert [c.shape for c in got] == [c.shape for c in ref], (y.R, [c.shape for c in ref])
    fr, fg = _full(ref), _full(got)
    assert np.linalg.norm(fg - fr) <= 1e-11 * np.linalg.norm(fr)
    if case == "redundant":
        assert list(y.R) == [1, 3, 4, 4, 2, 1]
        assert np.linalg.norm(fg - exact) <= 1e-12 * np.linalg.norm(exact)
    if case == "random":
        # the rounding error bound: ||x - round(x)|| <= eps ||x||
        assert np.linalg.norm(fg - exact) <= 1e-1 * np.linalg.norm(exact) * (1 + 1e-12)
    if case == "capped":
        assert max(y.R) <= 3


def test_round_operator_cores():
    """TT-matrix (4-D cores): L + L rounds to the rank of L, as TT.round does on the GPU."""
    import torchtt_amd as tt
    from torchtt_amd import workloads as W
    A, _ = W.laplace_potential(5, 8, seed=0)
    L = tt.TT([c.cuda() for c in A])
    S = (L + L)
    R = S.round(1e-12)
    assert all(c.is_cuda for c in R.cores)
    assert list(R.R) == list(L.R) or max(R.R) <= max(L.R)
    diff = (R - 2 * L).norm() / (2 * L).norm()
    assert float(diff) < 1e-12


def test_round_errors():
    import torchtt_amd as tt
    with pytest.raises(Exception):
        tt.solvers.round_tt("not a tt")
    x = tt.TT([torch.ones(1, 3, 1, dtype=torch.float64)])
    with pytest.raises(RuntimeError):
        tt.solvers.round_tt(x)  # CPU cores: the native path runs on the GPU only


@pytest.mark.parametrize("N,R", [([5, 4, 6, 3], [1, 3, 4, 2, 1]), ([7], [1, 1]), ([2, 3], [1, 5, 1]),
                                 ([8, 8, 8, 8, 8], [1, 6, 9, 9, 6, 1])])
def test_tt_full_matches_oracle(N, R):
    """Native full tensor of a TT vector (replacing torchttcpp.tt_full, cpp/full.h:3-12) against the
    oracle's contraction and the reference Python TT.full's einsum chain."""
    import torchtt_amd as tt
    rng = np.random.default_rng(len(N))
    cores = [rng.standard_normal((R[k], N[k], R[k + 1])) for k in range(len(N))]
    x = tt.TT([torch.tensor(c, device="cuda") for c in cores])
    print("tt_full", N, R, flush=True)
    full = tt.solvers.tt_full(x).cpu().numpy()
    ref = O.tt_full(cores)
    assert full.shape == tuple(N)
    assert np.linalg.norm(full - ref) <= 1e-13 * np.linalg.norm(ref)
    assert np.linalg.norm(full - x.full().cpu().numpy()) <= 1e-13 * np.linalg.norm(ref)
