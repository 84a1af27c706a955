"""Batched and served solves: each result equals the same system solved alone.

The solver's kernels are deterministic (fixed reduction orders, no atomics in the
arithmetic), so "the same" is checked bitwise; the stated tolerance would be 1e-12.
"""
import sys

import numpy as np
import pytest
import torch

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _small_systems(count, seed):
    import torchtt_amd as tt
    from torchtt_amd import workloads as W
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        d = int(rng.integers(3, 7))
        n = int(rng.integers(8, 21))
        A, b = W.anisotropic_diffusion(rng, d, n, int(rng.integers(2, 5)), int(rng.integers(1, 4)))
        out.append((tt.TT([c.cuda() for c in A]), tt.TT([c.cuda() for c in b]), int(rng.integers(20, 40))))
    return out


def _same(x, y):
    return x.R == y.R and all(torch.equal(a, b) for a, b in zip(x.cores, y.cores))


def _solo(systems, z0, **kw):
    import torchtt_amd as tt
    return [tt.solvers.amen_solve(A, b, rmax=rm, _z0=z, **kw) for (A, b, rm), z in zip(systems, z0)]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_equals_solo(precision):
    import torchtt_amd as tt
    systems = _small_systems(10, seed=3)
    torch.manual_seed(1)
    z0 = [tt.solvers._draw_z(s[1].N, 4, 0, "cpu") for s in systems]
    _say("batch vs solo", precision)
    xs = tt.solvers.amen_solve_batch(systems, eps=1e-10, precision=precision, workers=4, z0=z0)
    solo = _solo(systems, z0, eps=1e-10, precision=precision)
    for i, (x, y) in enumerate(zip(xs, solo)):
        assert _same(x, y), i
        assert tt.solvers.residual(systems[i][0], x, systems[i][1]) < 1e-8, i


def test_serve_equals_solo_and_admission():
    import torchtt_amd as tt
    systems = _small_systems(9, seed=4)
    torch.manual_seed(2)
    z0 = [tt.solvers._draw_z(s[1].N, 4, 0, "cpu") for s in systems]
    arrivals = [0.004 * i for i in range(len(systems))]
    _say("serve: 3 slots, overlapping arrivals")
    xs, info = tt.serving.serve(systems, arrivals, slots=3, eps=1e-10, z0=z0)
    solo = _solo(systems, z0, eps=1e-10)
    for i, (x, y) in enumerate(zip(xs, solo)):
        assert _same(x, y), i
    st, dn = info["start_s"], info["done_s"]
    assert np.all(st >= np.array(arrivals) - 1e-6)
    assert np.all(dn >= st)
    assert np.all(np.diff(st) >= -1e-9), "admission is in arrival order"
    # overlap happened: some request started before an earlier one finished
    assert any(st[i + 1] < dn[i] for i in range(len(systems) - 1))
    _say("  latencies ms", np.round(info["latency_s"] * 1e3, 1))

    # memory for one request at a time: no two requests overlap
    lib = tt._lib.load()
    need = []
    for (A, b, rm) in systems:
        N = tt._lib.i64_array(b.N)
        need.append(lib.ttamen_amen_solve_workspace_size(len(b.N), N, tt._lib.i64_array(A.R),
                                                         tt._lib.i64_array(b.R), None, rm, 4, 0, 40, 500, 0))
    mem = int(max(need)) + 4096
    _say("serve: memory for one request (%d bytes)" % mem)
    xs2, info2 = tt.serving.serve(systems, arrivals, slots=3, memory_bytes=mem, eps=1e-10, z0=z0)
    for i, (x, y) in enumerate(zip(xs2, solo)):
        assert _same(x, y), i
    st2, dn2 = info2["start_s"], info2["done_s"]
    assert all(st2[i + 1] >= dn2[i] - 1e-6 for i in range(len(systems) - 1))
