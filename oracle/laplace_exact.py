"""Closed-form solution of the d-dimensional discrete Laplace equation -- TEST INFRASTRUCTURE.

Solves (sum_k I x ... x T x ... x I) x = b for T = tridiag(-1, 2, -1) / h^2 (n x n,
Dirichlet) and a rank-1 right-hand side b = b_1 x ... x b_d by fast diagonalisation:
T = V diag(lam) V^T with the sine eigenvectors V[i, j] = sqrt(2/(n+1)) sin((i+1)(j+1) pi/(n+1))
and lam_j = (2 - 2 cos((j+1) pi / (n+1))) / h^2, hence

    x = (V x ... x V) [ (V^T b_1) x ... x (V^T b_d) / (lam_{j1} + ... + lam_{jd}) ].

This does not use the reference at all; it pins the AMEn solution for the Laplacian
workload (d = 6, n = 16, h = 1/17, b = ones) independently of any AMEn code.
"""
import numpy as np


def sine_basis(n, h):
    i = np.arange(1, n + 1)
    V = np.sqrt(2.0 / (n + 1)) * np.sin(np.outer(i, i) * np.pi / (n + 1))
    lam = (2.0 - 2.0 * np.cos(i * np.pi / (n + 1))) / h ** 2
    return V, lam


def laplace_solve_full(d, n, h, b_factors=None):
    """Dense solution tensor (n,)*d. b_factors: list of d vectors (default all ones)."""
    V, lam = sine_basis(n, h)
    if b_factors is None:
        b_factors = [np.ones(n)] * d
    # eigen-coefficients of the rhs and the denominator grid
    coef = None
    den = None
    for k in range(d):
        c = V.T @ b_factors[k]
        coef = c if coef is None else np.multiply.outer(coef, c)
        den = lam if den is None else np.add.outer(den, lam)
    y = coef / den
    # back-transform mode by mode
    for k in range(d):
        y = np.tensordot(V, y, axes=([1], [k]))
        y = np.moveaxis(y, 0, k)
    return y


def laplacian_dense_1d(n, h):
    return (2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)) / h ** 2
