"""Tensor-train container used at the solver's boundary.

Mirrors the parts of the reference's `torchtt.TT` (torchtt/_tt_base.py) that callers of
`amen_solve` touch: construction from cores or from a full tensor, the shape/rank
properties, `+ - * @ **`, `norm`, `round`, `full` and device moves. The arithmetic here
is torch on whatever device the cores live on; it builds inputs and checks results
(e.g. the global residual ||A x - b|| / ||b||) and is not part of the solve itself.
"""
import sys
import numpy as np
import torch

from .errors import ShapeMismatch, RankMismatch, IncompatibleTypes, InvalidArguments


def _rank_chop(s, eps):
    """Smallest r with ||s[r:]|| < eps (reference cpp/ortho.h:16-44)."""
    n = s.shape[0]
    if float(torch.linalg.norm(s)) == 0.0:
        return 1
    if eps <= 0.0:
        return n
    tail = torch.flip(torch.cumsum(torch.flip(s * s, [0]), 0), [0])  # tail[r] = sum s[r:]^2
    r = n - 1
    while r > 0 and float(tail[r]) < eps * eps:
        r -= 1
    return max(r + 1, 1)


def _tt_svd(full, shape, eps, rmax):
    """TT decomposition of a full tensor by sequential truncated SVDs."""
    d = len(shape)
    if d == 1:
        return [full.reshape(1, shape[0], 1)]
    delta = eps / np.sqrt(d - 1) * float(torch.linalg.norm(full))
    cores = []
    cur = full.reshape(shape[0], -1)
    r = 1
    for k in range(d - 1):
        cur = cur.reshape(r * shape[k], -1)
        u, s, vh = torch.linalg.svd(cur, full_matrices=False)
        rk = _rank_chop(s, delta)
        rk = min(rk, rmax if isinstance(rmax, int) else rmax[k + 1])
        cores.append(u[:, :rk].reshape(r, shape[k], rk))
        cur = s[:rk, None] * vh[:rk, :]
        r = rk
    cores.append(cur.reshape(r, shape[-1], 1))
    return cores


class TT:
    """A tensor (cores r x n x r') or operator (cores r x m x n x r') in TT format."""

    def __init__(self, source, shape=None, eps=1e-10, rmax=sys.maxsize):
        if source is None:
            self.cores, self._N, self._M, self._R, self._is_ttm = [], [], [], [1, 1], False
            return
        if isinstance(source, list):
            R = [source[0].shape[0]]
            N, M = [], []
            for i, c in enumerate(source):
                if c.shape[0] != R[-1]:
                    raise RankMismatch(
                        "Ranks of the given cores do not match: for core number %d previous rank "
                        "is %d and current rank is %d." % (i, R[-1], c.shape[0]))
                if c.dim() == 3:
                    N.append(c.shape[1]); R.append(c.shape[2])
                elif c.dim() == 4:
                    M.append(c.shape[1]); N.append(c.shape[2]); R.append(c.shape[3])
                else:
                    raise InvalidArguments("Invalid input: TT-cores have to be either 4d or 3d.")
            if R[0] != 1 or R[-1] != 1 or (len(M) != 0 and len(M) != len(N)):
                raise InvalidArguments("Check the ranks and the mode size.")
            self.cores = source
            self._R, self._N = R, N
            self._is_ttm = len(M) == len(N) and len(M) > 0
            self._M = M if self._is_ttm else []
            return
        if isinstance(source, np.ndarray):
            source = torch.from_numpy(source)
        if not torch.is_tensor(source):
            raise NotImplementedError("TT: source must be a list of cores, a torch tensor or a numpy array.")
        if shape is not None and isinstance(shape, list) and len(shape) > 0 and isinstance(shape[0], tuple):
            M = [s[0] for s in shape]
            N = [s[1] for s in shape]
            d = len(shape)
            t = source.reshape(M + N)
            perm = [v for k in range(d) for v in (k, d + k)]
            t = t.permute(perm).reshape([M[k] * N[k] for k in range(d)])
            cores = _tt_svd(t, [M[k] * N[k] for k in range(d)], eps, rmax)
            self.cores = [c.reshape(c.shape[0], M[k], N[k], c.shape[2]) for k, c in enumerate(cores)]
            self._M, self._N, self._is_ttm = M, N, True
        else:
            N = list(source.shape) if shape is None else list(shape)
            self.cores = _tt_svd(source.reshape(N), N, eps, rmax)
            self._M, self._N, self._is_ttm = [], N, False
        self._R = [1] + [c.shape[-1] for c in self.cores]

    # ------------------------------------------------------------------ properties
    @property
    def is_ttm(self):
        return self._is_ttm

    @property
    def N(self):
        return list(self._N)

    @property
    def M(self):
        if not self._is_ttm:
            raise IncompatibleTypes("The field is_ttm is defined only for TT matrices.")
        return list(self._M)

    @property
    def R(self):
        return list(self._R)

    @property
    def shape(self):
        return [(m, n) for m, n in zip(self._M, self._N)] if self._is_ttm else list(self._N)

    def __repr__(self):
        kind = "TT-matrix" if self._is_ttm else "TT"
        dev = self.cores[0].device if self.cores else "-"
        dt = self.cores[0].dtype if self.cores else "-"
        return "%s with sizes and ranks:\nN = %s\n%sR = %s\n\nDevice: %s, dtype: %s" % (
            kind, self._N, ("M = %s\n" % self._M) if self._is_ttm else "", self._R, dev, dt)

    # ------------------------------------------------------------------ conversions
    def clone(self):
        return TT([c.clone() for c in self.cores])

    def to(self, device=None, dtype=None):
        return TT([c.to(device=device, dtype=dtype) for c in self.cores])

    def cuda(self, device=None):
        return TT([c.cuda(device) for c in self.cores])

    def cpu(self):
        return TT([c.cpu() for c in self.cores])

    def full(self):
        t = self.cores[0]
        for c in self.cores[1:]:
            t = torch.tensordot(t, c, ([t.dim() - 1], [0]))
        t = t.reshape(t.shape[1:-1])
        if self._is_ttm:
            d = len(self._N)
            perm = [2 * k for k in range(d)] + [2 * k + 1 for k in range(d)]
            t = t.permute(perm)
        return t

    def numpy(self):
        return self.full().cpu().numpy()

    # ------------------------------------------------------------------ algebra
    def _check_same(self, other):
        if not isinstance(other, TT):
            raise InvalidArguments("Second operand must be a TT instance.")
        if self._is_ttm != other._is_ttm:
            raise IncompatibleTypes("Both operands must be TT tensors or TT matrices.")
        if self._N != other._N or (self._is_ttm and self._M != other._M):
            raise ShapeMismatch("Shapes are incompatible.")

    def __add__(self, other):
        if isinstance(other, (int, float)):
            return self + other * _ones_like(self)
        self._check_same(other)
        d = len(self._N)
        cores = []
        for k in range(d):
            a, b = self.cores[k], other.cores[k]
            if d == 1:
                cores.append(a + b)
            elif k == 0:
                cores.append(torch.cat([a, b], dim=-1))
            elif k == d - 1:
                cores.append(torch.cat([a, b], dim=0))
            else:
                ra, ra2, rb, rb2 = a.shape[0], a.shape[-1], b.shape[0], b.shape[-1]
                c = torch.zeros((ra + rb,) + tuple(a.shape[1:-1]) + (ra2 + rb2,), dtype=a.dtype, device=a.device)
                c[:ra, ..., :ra2] = a
                c[ra:, ..., ra2:] = b
                cores.append(c)
        return TT(cores)

    __radd__ = __add__

    def __neg__(self):
        return TT([-self.cores[0]] + list(self.cores[1:]))

    def __sub__(self, other):
        if isinstance(other, (int, float)):
            return self + (-other)
        return self + (-other)

    def __rsub__(self, other):
        return (-self) + other

    def __mul__(self, other):
        if isinstance(other, (int, float)) or (torch.is_tensor(other) and other.numel() == 1):
            return TT([self.cores[0] * other] + list(self.cores[1:]))
        self._check_same(other)
        cores = []
        for a, b in zip(self.cores, other.cores):
            if self._is_ttm:
                c = torch.einsum("aijb,cijd->acijbd", a, b)
                cores.append(c.reshape(a.shape[0] * b.shape[0], a.shape[1], a.shape[2], a.shape[3] * b.shape[3]))
            else:
                c = torch.einsum("aib,cid->acibd", a, b)
                cores.append(c.reshape(a.shape[0] * b.shape[0], a.shape[1], a.shape[2] * b.shape[2]))
        return TT(cores)

    __rmul__ = __mul__

    def __truediv__(self, other):
        """x / y: by a scalar, or elementwise by a TT through AMEn on the GPU with the reference's
        fixed arguments (torchtt/_tt_base.py:898-934: 50 sweeps, eps 1e-12, rank cap 500)."""
        if isinstance(other, (int, float)) or (torch.is_tensor(other) and other.numel() == 1):
            return self * (1.0 / other)
        if isinstance(other, TT):
            if self._is_ttm != other.is_ttm:
                raise IncompatibleTypes('Operands should be either TT or TTM.')
            if list(self.N) != list(other.N) or (self._is_ttm and list(self.M) != list(other.M)):
                raise ShapeMismatch('Both operands should have the same shape.')
            from .solvers import amen_divide
            return TT(amen_divide(other, self, 50, None, 1e-12, 500, verbose=False))
        raise InvalidArguments('Operand not permitted. A TT-object can be divided only with scalars.')

    def __rtruediv__(self, other):
        """scalar / x through AMEn on the GPU (torchtt/_tt_base.py:936-970)."""
        if isinstance(other, (int, float)) or (torch.is_tensor(other) and other.numel() == 1):
            o = TT([torch.ones((1, n, 1), dtype=self.cores[0].dtype, device=self.cores[0].device) for n in self.N])
            o.cores[0] *= other
            from .solvers import amen_divide
            return TT(amen_divide(self, o, 50, None, 1e-12, 500, verbose=False))
        raise InvalidArguments('The first operand must be int, float or 1d torch.tensor.')

    def __matmul__(self, other):
        if not (self._is_ttm and isinstance(other, TT)):
            raise InvalidArguments("The left operand of @ must be a TT-matrix and the right a TT.")
        if self._N != (other._M if other._is_ttm else other._N):
            raise ShapeMismatch("Shapes are incompatible.")
        cores = []
        for a, b in zip(self.cores, other.cores):
            if other._is_ttm:
                c = torch.einsum("amkb,cknd->acmnbd", a, b)
                cores.append(c.reshape(a.shape[0] * b.shape[0], a.shape[1], b.shape[2], a.shape[3] * b.shape[3]))
            else:
                c = torch.einsum("amkb,ckd->acmbd", a, b)
                cores.append(c.reshape(a.shape[0] * b.shape[0], a.shape[1], a.shape[3] * b.shape[2]))
        return TT(cores)

    def fast_matvec(self, other, eps=1e-12, initial=None, nswp=20, verb=False, use_cpp=True):
        """A @ x by DMRG sweeps on the GPU (reference torchtt/_tt_base.py:847-873)."""
        if not isinstance(other, TT):
            raise InvalidArguments('Second operand has to be TT object.')
        if not self._is_ttm or other.is_ttm:
            raise IncompatibleTypes('First operand should be a TT matrix and second a TT vector.')
        from ._dmrg import dmrg_matvec
        return dmrg_matvec(self, other, y0=initial, eps=eps, verb=verb, nswp=nswp, use_cpp=use_cpp)

    def __pow__(self, other):
        """Kronecker product (reference _tt_base.py:1341)."""
        if not isinstance(other, TT) or self._is_ttm != other._is_ttm:
            raise IncompatibleTypes("Kronecker product needs two TTs of the same kind.")
        return TT(list(self.cores) + list(other.cores))

    def t(self):
        if not self._is_ttm:
            raise IncompatibleTypes("Transpose is defined for TT-matrices.")
        return TT([c.permute(0, 2, 1, 3).contiguous() for c in self.cores])

    def norm(self, squared=False):
        """Frobenius norm via a left-to-right QR sweep (no squaring of small values)."""
        cur = self.cores[0]
        for k in range(len(self.cores) - 1):
            mode = list(cur.shape[1:-1])
            q, r = torch.linalg.qr(cur.reshape(-1, cur.shape[-1]))
            nxt = self.cores[k + 1]
            cur = (r @ nxt.reshape(nxt.shape[0], -1)).reshape([q.shape[1]] + list(nxt.shape[1:]))
            del mode
        nrm = torch.linalg.norm(cur)
        return nrm ** 2 if squared else nrm

    def round(self, eps=1e-12, rmax=sys.maxsize):
        """TT rounding: left-to-right QR, then right-to-left truncated SVD (rank_chop).

        FP64 cores on a ROCm GPU go through libttamen (solvers.round_tt, the reference's native
        round_this); other inputs use torch's QR/SVD with the same rule."""
        d = len(self.cores)
        if d <= 1:
            return self.clone()
        if self.cores[0].is_cuda and self.cores[0].dtype == torch.float64 and isinstance(rmax, int) \
                and max(self.R) <= 140:
            from .solvers import round_tt
            return round_tt(self, eps, rmax)
        cores = [c.clone() for c in self.cores]
        for k in range(d - 1):
            shp = cores[k].shape
            q, r = torch.linalg.qr(cores[k].reshape(-1, shp[-1]))
            cores[k] = q.reshape(tuple(shp[:-1]) + (q.shape[1],))
            nxt = cores[k + 1]
            cores[k + 1] = (r @ nxt.reshape(nxt.shape[0], -1)).reshape((q.shape[1],) + tuple(nxt.shape[1:]))
        delta = eps / np.sqrt(d - 1)
        for k in range(d - 1, 0, -1):
            shp = cores[k].shape
            u, s, vh = torch.linalg.svd(cores[k].reshape(shp[0], -1), full_matrices=False)
            rk = _rank_chop(s, delta * float(torch.linalg.norm(s)))
            rk = min(rk, rmax if isinstance(rmax, int) else rmax[k])
            cores[k] = vh[:rk, :].reshape((rk,) + tuple(shp[1:]))
            prv = cores[k - 1]
            cores[k - 1] = (prv.reshape(-1, prv.shape[-1]) @ (u[:, :rk] * s[:rk])).reshape(tuple(prv.shape[:-1]) + (rk,))
        return TT(cores)


def _ones_like(x):
    if x.is_ttm:
        cores = [torch.ones((1, m, n, 1), dtype=x.cores[0].dtype, device=x.cores[0].device) for m, n in zip(x.M, x.N)]
    else:
        cores = [torch.ones((1, n, 1), dtype=x.cores[0].dtype, device=x.cores[0].device) for n in x.N]
    return TT(cores)
