"""Synthetic AMEn workloads (operator and right-hand-side TT cores).

Each builder returns plain torch FP64 CPU cores: operator cores rA x n x n x rA',
right-hand-side cores rb x n x rb'. They define the solver benchmarks:

  laplace_ones(d, n)         d-dim Dirichlet Laplacian, stencil [-1, 2, -1] / h^2 with
                             h = 1/(n+1), canonical TT rank 2; b = all ones (rank 1).
                             (d=6, n=16: the reference's CPU-runnable case.)
  laplace_potential(d, n)    Laplacian + diag(1 + c_i x^2) per mode, c_i ~ U(0.5, 2),
                             x_j = j h; the two rank-2 operators are block-concatenated
                             (TT rank 4, not rounded); b = random rank-4 TT with N(0,1)
                             cores scaled to ||b|| = 1.  (d=10, n=64 and d=12, n=64.)
  anisotropic_diffusion(...) sum_i a_i T_i with a_i log-uniform in [0.1, 10]; optionally
                             + potential (rank 3-4); random rank-1..4 rhs, ||b|| = 1.
"""
import numpy as np
import torch


def laplacian_1d(n, h=None):
    h = 1.0 / (n + 1) if h is None else h
    T = 2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    return T / h ** 2


def sum_operator_cores(mats):
    """Canonical rank-2 TT cores of sum_k I x ... x M_k x ... x I (one matrix per mode)."""
    d = len(mats)
    cores = []
    for k, M in enumerate(mats):
        n = M.shape[0]
        I = np.eye(n)
        if d == 1:
            c = M.reshape(1, n, n, 1).copy()
        elif k == 0:
            c = np.zeros((1, n, n, 2)); c[0, :, :, 0] = M; c[0, :, :, 1] = I
        elif k == d - 1:
            c = np.zeros((2, n, n, 1)); c[0, :, :, 0] = I; c[1, :, :, 0] = M
        else:
            c = np.zeros((2, n, n, 2)); c[0, :, :, 0] = I; c[1, :, :, 0] = M; c[1, :, :, 1] = I
        cores.append(c)
    return cores


def block_concat_operator(a_cores, b_cores):
    """Cores of A + B by TT block concatenation (no rounding): ranks add."""
    d = len(a_cores)
    out = []
    for k in range(d):
        a, b = a_cores[k], b_cores[k]
        ra, m, n, ra2 = a.shape
        rb, _, _, rb2 = b.shape
        if d == 1:
            out.append(a + b)
        elif k == 0:
            out.append(np.concatenate([a, b], axis=3))
        elif k == d - 1:
            out.append(np.concatenate([a, b], axis=0))
        else:
            c = np.zeros((ra + rb, m, n, ra2 + rb2))
            c[:ra, :, :, :ra2] = a
            c[ra:, :, :, ra2:] = b
            out.append(c)
    return out


def tt_norm_np(cores):
    cur = cores[0]
    for k in range(len(cores) - 1):
        r0, n, r1 = cur.shape
        q, r = np.linalg.qr(cur.reshape(r0 * n, r1))
        nxt = cores[k + 1]
        cur = (r @ nxt.reshape(nxt.shape[0], -1)).reshape(q.shape[1], nxt.shape[1], nxt.shape[2])
    return float(np.linalg.norm(cur))


def random_rhs(rng, d, n_list, rank):
    r = [1] + [rank] * (d - 1) + [1]
    cores = [rng.standard_normal((r[k], n_list[k], r[k + 1])) for k in range(d)]
    nrm = tt_norm_np(cores)
    cores[0] = cores[0] / nrm
    return cores


def _to_torch(cores):
    return [torch.from_numpy(np.ascontiguousarray(c)) for c in cores]


def laplace_ones(d=6, n=16):
    A = sum_operator_cores([laplacian_1d(n)] * d)
    b = [np.ones((1, n, 1)) for _ in range(d)]
    return _to_torch(A), _to_torch(b)


def laplace_potential(d=10, n=64, seed=0, rhs_rank=4):
    rng = np.random.default_rng(seed)
    h = 1.0 / (n + 1)
    xg = np.arange(1, n + 1) * h
    c = rng.uniform(0.5, 2.0, size=d)
    L = sum_operator_cores([laplacian_1d(n, h)] * d)
    V = sum_operator_cores([np.diag(1.0 + c[i] * xg ** 2) for i in range(d)])
    A = block_concat_operator(L, V)
    b = random_rhs(rng, d, [n] * d, rhs_rank)
    return _to_torch(A), _to_torch(b)


def anisotropic_diffusion(rng, d, n, op_rank, rhs_rank):
    """One system of the batched workload: sum_i a_i T_i (+ potential when op_rank > 2)."""
    h = 1.0 / (n + 1)
    a = np.exp(rng.uniform(np.log(0.1), np.log(10.0), size=d))
    A = sum_operator_cores([a[i] * laplacian_1d(n, h) for i in range(d)])
    if op_rank > 2:
        xg = np.arange(1, n + 1) * h
        c = rng.uniform(0.5, 2.0, size=d)
        V = sum_operator_cores([np.diag(1.0 + c[i] * xg ** 2) for i in range(d)])
        if op_rank == 3:
            # rank-3 form: share the identity channel of the two sums
            A = _sum_rank3(A, V)
        else:
            A = block_concat_operator(A, V)
    b = random_rhs(rng, d, [n] * d, rhs_rank)
    return _to_torch(A), _to_torch(b)


def _sum_rank3(L, V):
    """Rank-3 cores of L + V for two canonical rank-2 sums over the same modes."""
    d = len(L)
    out = []
    for k in range(d):
        T, P = None, None
        n = L[k].shape[1]
        I = np.eye(n)
        if k == 0:
            T, P = L[k][0, :, :, 0], V[k][0, :, :, 0]
            c = np.zeros((1, n, n, 3)); c[0, :, :, 0] = T; c[0, :, :, 1] = P; c[0, :, :, 2] = I
        elif k == d - 1:
            T, P = L[k][1, :, :, 0], V[k][1, :, :, 0]
            c = np.zeros((3, n, n, 1)); c[0, :, :, 0] = I; c[1, :, :, 0] = I; c[2, :, :, 0] = T + P
        else:
            T, P = L[k][1, :, :, 0], V[k][1, :, :, 0]
            c = np.zeros((3, n, n, 3))
            c[0, :, :, 0] = I; c[1, :, :, 1] = I; c[2, :, :, 0] = T; c[2, :, :, 1] = P
            c[2, :, :, 2] = I
        out.append(c)
    return out


def batch_systems(count=256, seed=0):
    """The batched workload: uneven d, n, operator and rhs ranks, rank caps."""
    rng = np.random.default_rng(seed)
    systems = []
    for _ in range(count):
        d = int(rng.integers(4, 13))
        n = int(rng.integers(8, 65))
        op_rank = int(rng.integers(2, 5))
        rhs_rank = int(rng.integers(1, 5))
        rmax = int(rng.integers(16, 65))
        A, b = anisotropic_diffusion(rng, d, n, op_rank, rhs_rank)
        systems.append({"A": A, "b": b, "rmax": rmax, "d": d, "n": n})
    return systems
