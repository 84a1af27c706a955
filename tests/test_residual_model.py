"""CPU model of the native global residual (torchtt_amd/csrc/residual.hip): the same sweep, buffer
layout and strided GEMM descriptors (each emulated element by element from its strides), the same
Householder R factor, against the oracle's TT residual (QR-based norm of the assembled residual TT)
and against an exact cancellation case (b = A x up to a relative 1e-9 perturbation)."""
import numpy as np

import amen_oracle as O


def gemm(M, N, K, B1, B2, a, b, c, A, Bm, C, alpha=1.0):
    """Strided batched GEMM exactly as GemmDesc describes (flat arrays)."""
    for i1 in range(B1):
        for i2 in range(B2):
            for i in range(M):
                for j in range(N):
                    s = 0.0
                    for k in range(K):
                        s += A[a[0] + i1*a[3] + i2*a[4] + i*a[1] + k*a[2]] * Bm[b[0] + i1*b[3] + i2*b[4] + k*b[1] + j*b[2]]
                    C[c[0] + i1*c[3] + i2*c[4] + i*c[1] + j*c[2]] = alpha * s

def house_r(M, rows, n):
    Mc = M[:rows*n].reshape(n, rows).T.copy()  # column-major -> array
    k = min(rows, n)
    R = np.full((k, n), np.nan)  # only what the kernel writes is defined
    for j in range(k):
        x = Mc[j:, j].copy(); nrm = np.sqrt(x @ x); x0 = x[0]
        alpha = -nrm if x0 >= 0 else nrm
        R[j, :j] = 0.0
        R[j, j] = alpha if nrm > 0 else 0.0
        if nrm == 0:
            R[j, j+1:] = Mc[j, j+1:]; continue
        tau = 1.0 / (nrm*nrm - alpha*x0)
        v = x.copy(); v[0] = x0 - alpha
        for c in range(j+1, n):
            f = tau * (v @ Mc[j:, c]); Mc[j:, c] -= f * v; R[j, c] = Mc[j, c]
    return R, k

def residual(A, x, b, with_ax=True):
    d = len(b); N = [c.shape[1] for c in b]
    rA = [1] + [c.shape[3] for c in A]; rx = [1] + [c.shape[2] for c in x]; rb = [1] + [c.shape[2] for c in b]
    if not with_ax: rA = [0]*(d+1); rx = [0]*(d+1)
    Rp = np.array([1.0, -1.0]) if with_ax else np.array([1.0]); q = 1
    for k in range(d):
        n, ra, ra1, rxk, rx1, rbk, rb1 = N[k], rA[k], rA[k+1], rx[k], rx[k+1], rb[k], rb[k+1]
        L = ra*rxk + rbk; Lax1 = ra1*rx1; Lr = Lax1 + rb1; rows = q*n
        cur = np.zeros(rows*Lr)
        if with_ax:
            u = np.zeros(q*ra*n*rx1); xf = x[k].reshape(-1)
            gemm(q, n*rx1, rxk, 1, ra, (0, L, 1, 0, rxk), (0, n*rx1, 1, 0, 0), (0, ra*n*rx1, 1, 0, n*rx1), Rp, xf, u)
            Apm = A[k].transpose(1, 3, 0, 2).reshape(-1)  # [m][S][s][n]
            gemm(n, rx1, ra*n, q, ra1, (0, ra1*ra*n, 1, 0, ra*n), (0, rx1, 1, ra*n*rx1, 0), (0, 1, rows, n, rx1*rows), Apm, u, cur)
        gemm(q, n, rbk, 1, rb1, (ra*rxk, L, 1, 0, 0), (0, n*rb1, rb1, 0, 1), (Lax1*rows, n, 1, 0, rows), Rp, b[k].reshape(-1), cur)
        if k == d - 1: break
        R, q = house_r(cur, rows, Lr); Rp = R.reshape(-1)
    rows = q*N[d-1]
    v = cur[:rows] + cur[rows:2*rows] if with_ax else cur[:rows]
    return np.linalg.norm(v)

def test_residual_model_against_oracle():
    rng = np.random.default_rng(0)
    for d, n, ra, rxr, rbr in [(3, 3, 2, 2, 2), (1, 4, 1, 1, 1), (2, 3, 3, 2, 1), (4, 2, 2, 3, 2)]:
        ranks = lambda r: [1] + [r]*(d-1) + [1]
        RA, RX, RB = ranks(ra), ranks(rxr), ranks(rbr)
        A = [rng.standard_normal((RA[k], n, n, RA[k+1])) for k in range(d)]
        x = [rng.standard_normal((RX[k], n, RX[k+1])) for k in range(d)]
        b = [rng.standard_normal((RB[k], n, RB[k+1])) for k in range(d)]
        r = residual(A, x, b); nb = residual(A, x, b, False)
        ref = O.tt_residual(A, x, b)
        print(d, n, ra, rxr, rbr, "model %.12e ref %.12e  norm b %.6e vs %.6e" % (r / nb, ref, nb, O.tt_norm(b)))
        assert abs(r/nb - ref) <= 1e-10 * ref and abs(nb - O.tt_norm(b)) <= 1e-10 * nb
    # near-cancellation: b = A x exactly plus a tiny perturbation
    d, n = 3, 3
    A = [rng.standard_normal((1 if k == 0 else 2, n, n, 1 if k == d-1 else 2)) for k in range(d)]
    x = [rng.standard_normal((1 if k == 0 else 2, n, 1 if k == d-1 else 2)) for k in range(d)]
    Ax = [np.einsum("amnb,inj->aimbj", a, c).reshape(a.shape[0]*c.shape[0], n, a.shape[3]*c.shape[2]) for a, c in zip(A, x)]
    b = [c.copy() for c in Ax]; b[0] = b[0] * (1 + 1e-9)
    r = residual(A, x, b) / residual(A, x, b, False)
    print("cancellation: model %.6e exact %.6e" % (r, 1e-9 / (1 + 1e-9)))
    assert abs(r - 1e-9/(1+1e-9)) < 1e-3 * 1e-9
