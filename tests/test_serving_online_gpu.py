"""Online serving (torchtt_amd.serving.Server, libttamen ttamen_server_*): requests submitted while
others run are admitted as slots and memory free up, run together, and each result equals the
same request solved alone (bitwise: same kernels, same order, deterministic reductions)."""
import sys
import time

import numpy as np
import pytest
import torch

from conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _systems(count, seed):
    import torchtt_amd as tt
    from torchtt_amd import workloads as W
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        d, n = int(rng.integers(3, 7)), int(rng.integers(8, 21))
        A, b = W.anisotropic_diffusion(rng, d, n, int(rng.integers(2, 5)), int(rng.integers(1, 4)))
        out.append((tt.TT([c.cuda() for c in A]), tt.TT([c.cuda() for c in b]), int(rng.integers(20, 40))))
    return out


def _same(x, y):
    return x.R == y.R and all(torch.equal(a, b) for a, b in zip(x.cores, y.cores))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_online_server_submit_while_running(precision):
    import torchtt_amd as tt
    systems = _systems(8, seed=21)
    torch.manual_seed(5)
    z0 = [tt.solvers._draw_z(s[1].N, 4, 0, "cpu") for s in systems]
    _say("online server, 3 slots,", precision)
    with tt.serving.Server(slots=3, memory_bytes=1 << 30, eps=1e-10, precision=precision) as srv:
        reqs = [srv.submit(A, b, rmax=rm, z0=z) for (A, b, rm), z in zip(systems[:3], z0[:3])]
        t0 = time.time()
        while not any(r.state() >= 1 for r in reqs) and time.time() - t0 < 30:
            time.sleep(0.0005)
        for (A, b, rm), z in zip(systems[3:], z0[3:]):
            reqs.append(srv.submit(A, b, rmax=rm, z0=z))
            time.sleep(0.003)
        xs = [r.result() for r in reqs]
        times = np.array([r.times() for r in reqs])
    sub, start, done = times[:, 0], times[:, 1], times[:, 2]
    _say("  latency ms", np.round((done - sub) * 1e3, 1))
    assert np.all(start >= sub - 1e-9) and np.all(done >= start)
    assert np.all(np.diff(start) >= -1e-9), "admission is in submission order"
    # requests submitted while others ran were admitted before those finished (continuous batching)
    assert any(start[j] < done[i] for i in range(3) for j in range(3, len(reqs)))
    for i, ((A, b, rm), z) in enumerate(zip(systems, z0)):
        y = tt.solvers.amen_solve(A, b, rmax=rm, eps=1e-10, precision=precision, _z0=z)
        assert _same(xs[i], y), i


def test_online_server_failing_request():
    """A request that fails (its solution needs rank 144 at an interface, above the rank capacity
    served solves are planned at, 140: status 4) reports its status; requests around it are unaffected."""
    import torchtt_amd as tt
    g = torch.Generator().manual_seed(5)
    N, rb, r0 = [16] * 4, [1, 16, 200, 16, 1], [1, 16, 140, 16, 1]
    bad_b = tt.TT([torch.randn(rb[k], N[k], rb[k + 1], dtype=torch.float64, generator=g).cuda() for k in range(4)])
    bad_x0 = tt.TT([torch.randn(r0[k], N[k], r0[k + 1], dtype=torch.float64, generator=g).cuda() for k in range(4)])
    systems = _systems(2, seed=22)
    torch.manual_seed(6)
    z0 = [tt.solvers._draw_z(s[1].N, 4, 0, "cpu") for s in systems]
    with tt.serving.Server(slots=2, memory_bytes=1 << 30, eps=1e-13, nswp=2) as srv:
        r0_ = srv.submit(*systems[0][:2], rmax=systems[0][2], z0=z0[0])
        bad = srv.submit(tt.eye(N, device="cuda"), bad_b, x0=bad_x0)
        r1_ = srv.submit(*systems[1][:2], rmax=systems[1][2], z0=z0[1])
        with pytest.raises(tt._lib.TTAmenError) as ei:
            bad.result()
        assert ei.value.code == 4 and "not supported" in str(ei.value)
        x0_, x1_ = r0_.result(), r1_.result()
    for x, (A, b, rm), z in ((x0_, systems[0], z0[0]), (x1_, systems[1], z0[1])):
        y = tt.solvers.amen_solve(A, b, rmax=rm, eps=1e-13, nswp=2, _z0=z)
        assert _same(x, y)
