"""Test matrices for the dense QR / SVD kernels, shared by the GPU kernel tests and the CPU
model of the QR algorithm."""
import numpy as np


def qr_case(kind, m, n, seed):
    """An m x n FP64 matrix of the given kind, drawn from a generator seeded with `seed`."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n))
    if kind == "zero_cols":
        A[:, 1] = 0.0
        A[:, n - 1] = 0.0
    elif kind == "dependent":
        A[:, n - 1] = A[:, 0] * 2.0 - A[:, 1]
    elif kind == "graded":
        A = A * np.logspace(0, -14, n)[None, :]
    elif kind == "tiny_tail":
        # an orthonormal block plus a tiny residual block (the enrichment QR shape)
        q, _ = np.linalg.qr(rng.standard_normal((m, n - 4)))
        A = np.concatenate([q, 1e-11 * rng.standard_normal((m, 4))], 1)
    elif kind == "rank_deficient":
        # numerical rank 2/3 of the columns, singular values 1 .. 1e-10 then ~1e-19 (the shape
        # of an AMEn local solution on a near-square first core); a three-pass Cholesky-QR
        # leaves Q^T Q off by 1e-2 on such inputs
        k = min(m, n)
        r = max(1, (2 * k) // 3)
        u, _ = np.linalg.qr(rng.standard_normal((m, k)))
        v, _ = np.linalg.qr(rng.standard_normal((n, k)))
        s = np.concatenate([np.geomspace(1.0, 1e-10, r), 1e-19 * rng.random(k - r)])
        A = (u * s) @ v.T
    elif kind == "square_zero_col":
        # a zero column in a nearly square matrix: few unit vectors keep half their norm
        # after projecting out the other columns, so the completion must pick its row
        A[:, n // 2] = 0.0
    return A
