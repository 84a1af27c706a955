"""TT constructors used to build solver inputs (reference torchtt/_extras.py: eye 19,
zeros 41, kron 75, ones 113, random 271, randn 306, dot 477)."""
import numpy as np
import torch

from ._tt import TT
from .errors import InvalidArguments, ShapeMismatch


def _is_ttm_shape(N):
    return len(N) > 0 and isinstance(N[0], tuple)


def eye(shape, dtype=torch.float64, device=None):
    """Identity TT-matrix with mode sizes `shape` (rank 1)."""
    shape = list(shape)
    return TT([torch.eye(n, dtype=dtype, device=device).reshape(1, n, n, 1) for n in shape])


def ones(N, dtype=torch.float64, device=None):
    N = list(N)
    if _is_ttm_shape(N):
        return TT([torch.ones((1, m, n, 1), dtype=dtype, device=device) for m, n in N])
    return TT([torch.ones((1, n, 1), dtype=dtype, device=device) for n in N])


def zeros(N, dtype=torch.float64, device=None):
    N = list(N)
    if _is_ttm_shape(N):
        return TT([torch.zeros((1, m, n, 1), dtype=dtype, device=device) for m, n in N])
    return TT([torch.zeros((1, n, 1), dtype=dtype, device=device) for n in N])


def _ranks(R, d):
    if isinstance(R, int):
        return [1] + [R] * (d - 1) + [1]
    R = list(R)
    if len(R) != d + 1 or R[0] != 1 or R[-1] != 1:
        raise InvalidArguments("Check the ranks and the mode size.")
    return R


def random(N, R, dtype=torch.float64, device=None):
    """TT with independent N(0,1) core entries (reference _extras.py:271)."""
    N = list(N)
    d = len(N)
    R = _ranks(R, d)
    if _is_ttm_shape(N):
        return TT([torch.randn((R[k], N[k][0], N[k][1], R[k + 1]), dtype=dtype, device=device) for k in range(d)])
    return TT([torch.randn((R[k], N[k], R[k + 1]), dtype=dtype, device=device) for k in range(d)])


def randn(N, R, var=1.0, dtype=torch.float64, device=None):
    """Random TT whose full tensor has entries of variance `var` (cores scaled accordingly)."""
    N = list(N)
    d = len(N)
    R = _ranks(R, d)
    v1 = var / np.prod(R)
    v = v1 ** (1.0 / d)
    if _is_ttm_shape(N):
        return TT([torch.randn((R[k], N[k][0], N[k][1], R[k + 1]), dtype=dtype, device=device) * np.sqrt(v) for k in range(d)])
    return TT([torch.randn((R[k], N[k], R[k + 1]), dtype=dtype, device=device) * np.sqrt(v) for k in range(d)])


def kron(a, b):
    """Kronecker product of two TTs (cores concatenated)."""
    if a is None:
        return b
    if b is None:
        return a
    return a ** b


def dot(a, b):
    """Inner product <a, b> of two TT tensors."""
    if a.N != b.N:
        raise ShapeMismatch("Shapes are incompatible.")
    phi = torch.ones((1, 1), dtype=a.cores[0].dtype, device=a.cores[0].device)
    for ca, cb in zip(a.cores, b.cores):
        phi = torch.einsum("ab,aic,bid->cd", phi, ca, cb)
    return phi.squeeze()


def meshgrid(vectors):
    """Rank-1 TTs X_i with vectors[i] on mode i and ones elsewhere (reference
    torchtt/_extras.py:455-474)."""
    dtype = vectors[0].dtype
    out = []
    for i in range(len(vectors)):
        cores = [torch.ones((1, v.shape[0], 1), dtype=dtype, device=v.device) for v in vectors]
        cores[i] = vectors[i].reshape(1, -1, 1)
        out.append(TT(cores))
    return out


def elementwise_divide(x, y, eps=1e-12, starting_tensor=None, nswp=50, kick=4, local_iterations=40, resets=2,
                       preconditioner=None, verbose=False):
    """x / y elementwise by AMEn on the GPU with adjustable arguments (reference
    torchtt/_extras.py:588-612): amen_divide(y, x, ...) with rank cap 1000."""
    from .solvers import amen_divide
    if not isinstance(x, TT):
        o = TT([torch.ones((1, n, 1), dtype=y.cores[0].dtype, device=y.cores[0].device) for n in y.N])
        o.cores[0] *= x
        x = o
    cores = amen_divide(y, x, nswp, starting_tensor, eps, rmax=1000, kickrank=kick,
                        local_iterations=local_iterations, resets=resets, verbose=verbose,
                        preconditioner=preconditioner)
    return TT(cores)
