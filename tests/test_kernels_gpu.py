"""Kernel-level parity on the GPU: each C-ABI entry point against an FP64 CPU reference
(the oracle's contractions, numpy linear algebra, or a torch fp32/fp64 product)."""
import ctypes
import sys

import numpy as np
import pytest
import torch

import amen_oracle as O
from conftest import gpu_available
from qr_cases import qr_case

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a ROCm GPU")]

DEV = "cuda"


def _lib():
    from torchtt_amd import _lib as L
    return L, L.load()


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ws(nbytes):
    return torch.empty(int(nbytes), dtype=torch.uint8, device=DEV)


def _rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# ----------------------------------------------------------------- GEMM
@pytest.mark.parametrize("M,N,K,batch,trans", [(64, 64, 64, 1, "nn"), (100, 37, 260, 3, "tn"),
                                               (5, 7, 3, 2, "nt"), (84, 84, 5376, 1, "tn"),
                                               (300, 20, 1000, 1, "tt"), (1, 1, 1, 1, "nn"),
                                               (64, 48, 96, 2, "tn"), (64, 48, 96, 2, "nt"),
                                               (64, 48, 96, 2, "tt"), (130, 66, 34, 3, "nn"),
                                               (130, 67, 34, 2, "tt"), (129, 66, 35, 2, "nt"),
                                               # split K: 63 slices of 80 with a ragged last one of 40
                                               (40, 33, 5000, 2, "nt")])
def test_gemm_f64(M, N, K, batch, trans):
    L, lib = _lib()
    _say("gemm", M, N, K, batch, trans)
    g = torch.Generator().manual_seed(M * 7 + N)
    A = torch.randn(batch, M, K, dtype=torch.float64, generator=g)
    B = torch.randn(batch, K, N, dtype=torch.float64, generator=g)
    C0 = torch.randn(batch, M, N, dtype=torch.float64, generator=g)
    Ad = (A.transpose(1, 2).contiguous() if trans[0] == "t" else A).to(DEV)
    Bd = (B.transpose(1, 2).contiguous() if trans[1] == "t" else B).to(DEV)
    Cd = C0.clone().to(DEV)
    sa = (1, M) if trans[0] == "t" else (K, 1)
    sb = (1, K) if trans[1] == "t" else (N, 1)
    ws = _ws(64 << 20)
    code = lib.ttamen_gemm_f64(M, N, K, batch, 1.5, _p(Ad), sa[0], sa[1], M * K, _p(Bd), sb[0], sb[1], K * N,
                               -0.5, _p(Cd), N, 1, M * N, _p(ws), ws.numel(), _stream())
    L.check("gemm", code)
    torch.cuda.synchronize()
    want = 1.5 * torch.bmm(A, B) - 0.5 * C0
    assert _rel(Cd.cpu().numpy(), want.numpy()) < 1e-13


@pytest.mark.parametrize("trans", ["nn", "tn", "nt", "tt"])
def test_gemm_f64_unaligned_base(trans):
    """Operands starting one element past a 16-byte boundary (even sizes and strides): the tile
    must stage them element by element."""
    L, lib = _lib()
    M, N, K, batch = 64, 32, 48, 2
    _say("gemm unaligned base", trans)
    g = torch.Generator().manual_seed(5)
    A = torch.randn(batch, M, K, dtype=torch.float64, generator=g)
    B = torch.randn(batch, K, N, dtype=torch.float64, generator=g)
    Ab = torch.zeros(batch * M * K + 1, dtype=torch.float64, device=DEV)
    Bb = torch.zeros(batch * K * N + 1, dtype=torch.float64, device=DEV)
    Ab[1:] = (A.transpose(1, 2).contiguous() if trans[0] == "t" else A).reshape(-1).to(DEV)
    Bb[1:] = (B.transpose(1, 2).contiguous() if trans[1] == "t" else B).reshape(-1).to(DEV)
    Cd = torch.zeros(batch, M, N, dtype=torch.float64, device=DEV)
    sa = (1, M) if trans[0] == "t" else (K, 1)
    sb = (1, K) if trans[1] == "t" else (N, 1)
    ws = _ws(16 << 20)
    pa = Ab.data_ptr() + 8
    pb = Bb.data_ptr() + 8
    code = lib.ttamen_gemm_f64(M, N, K, batch, 1.0, ctypes.c_void_p(pa), sa[0], sa[1], M * K,
                               ctypes.c_void_p(pb), sb[0], sb[1], K * N, 0.0, _p(Cd), N, 1, M * N,
                               _p(ws), ws.numel(), _stream())
    L.check("gemm", code)
    torch.cuda.synchronize()
    assert _rel(Cd.cpu().numpy(), torch.bmm(A, B).numpy()) < 1e-13


# ----------------------------------------------------------------- local operator
SHAPES = [(84, 4, 84, 64, 4, 84, 84, 1), (7, 2, 5, 16, 2, 6, 4, 3), (4, 4, 20, 64, 4, 4, 24, 1),
          (1, 1, 1, 16, 2, 5, 5, 2), (20, 2, 20, 8, 1, 1, 1, 1),
          # fused local operator: ragged l and R blocks, a rank-3 operator, 17 partial sums of R,
          # a residual-proxy product, and a long mode whose FP64 tiles exceed the LDS (three GEMMs)
          (13, 3, 11, 17, 3, 9, 23, 1), (132, 4, 132, 64, 4, 132, 132, 1), (30, 4, 30, 33, 4, 4, 29, 2),
          (5, 4, 6, 160, 4, 3, 8, 1)]


def _rand_local(shape, dtype=torch.float64, seed=0):
    l1, s, r1, n, S, l2, r2, nv = shape
    g = torch.Generator().manual_seed(seed)
    pl = torch.randn(l1, s, r1, dtype=dtype, generator=g)
    a = torch.randn(s, n, n, S, dtype=dtype, generator=g)
    pr = torch.randn(l2, S, r2, dtype=dtype, generator=g)
    x = torch.randn(nv, r1, n, r2, dtype=dtype, generator=g)
    return pl, a, pr, x


@pytest.fixture(params=[0, 1, 2], ids=["gemm+tail", "fused", "gemms"])
def matvec_mode(request):
    """The three forms of the local operator: a GEMM and the fused stages 2-3 kernel (default), the
    one-launch kernel, and three GEMMs."""
    L, lib = _lib()
    prev = lib.ttamen_matvec_mode(request.param)
    yield request.param
    lib.ttamen_matvec_mode(prev)


@pytest.mark.parametrize("shape", SHAPES)
def test_local_matvec_f64(shape, matvec_mode):
    L, lib = _lib()
    l1, s, r1, n, S, l2, r2, nv = shape
    _say("local_matvec f64", shape, "mode", matvec_mode)
    pl, a, pr, x = _rand_local(shape)
    y = torch.empty(nv, l1, n, l2, dtype=torch.float64, device=DEV)
    ws = _ws(lib.ttamen_local_matvec_workspace_size(l1, s, r1, n, S, l2, r2, nv))
    t = [v.to(DEV) for v in (pl, a, pr, x)]
    code = lib.ttamen_local_matvec(_p(t[0]), _p(t[1]), _p(t[2]), _p(t[3]), _p(y), l1, s, r1, n, S, l2, r2, nv,
                                   _p(ws), ws.numel(), _stream())
    L.check("local_matvec", code)
    torch.cuda.synchronize()
    for v in range(nv):
        want = O.local_product(pr.numpy(), pl.numpy(), a.numpy(), x[v].numpy())
        assert _rel(y[v].cpu().numpy(), want) < 1e-13, v


@pytest.mark.parametrize("shape", SHAPES)
def test_local_matvec_f32(shape, matvec_mode):
    L, lib = _lib()
    l1, s, r1, n, S, l2, r2, nv = shape
    _say("local_matvec f32", shape, "mode", matvec_mode)
    pl, a, pr, x = _rand_local(shape, torch.float32, seed=1)
    y = torch.empty(nv, l1, n, l2, dtype=torch.float32, device=DEV)
    ws = _ws(lib.ttamen_local_matvec_workspace_size(l1, s, r1, n, S, l2, r2, nv) * 2)
    t = [v.to(DEV) for v in (pl, a, pr, x)]
    code = lib.ttamen_local_matvec_f32(_p(t[0]), _p(t[1]), _p(t[2]), _p(t[3]), _p(y), l1, s, r1, n, S, l2, r2,
                                       nv, _p(ws), ws.numel(), _stream())
    L.check("local_matvec_f32", code)
    torch.cuda.synchronize()
    for v in range(nv):
        want = O.local_product(pr.double().numpy(), pl.double().numpy(), a.double().numpy(), x[v].double().numpy())
        # FP32 accumulation over K = r1, s*n, S*r2 terms: relative error ~ sqrt(K) * 2^-24
        assert _rel(y[v].cpu().double().numpy(), want) < 2e-5, v


def _block_sparse_core(s, n, S, seed):
    """An operator core with the block structure of a TT Laplacian-plus-potential: identities, a
    tridiagonal stencil and a diagonal on some (s, S) blocks, zero elsewhere; with S = 5 the last S
    value has no nonzero at all, and with s = 4 one entry sits alone in the last row and column."""
    g = np.random.default_rng(seed)
    a = np.zeros((s, n, n, S))
    eye = np.eye(n)
    tri = np.diag(g.standard_normal(n)) + np.diag(g.standard_normal(n - 1), 1) + np.diag(g.standard_normal(n - 1), -1)
    a[0, :, :, 0] = eye
    if s > 1:
        a[1, :, :, 0] = tri
        a[1, :, :, 1 % S] += eye
    if s > 3 and S > 3:
        a[2, :, :, 2] = eye
        a[3, :, :, 2] = np.diag(g.standard_normal(n))
        a[3, n - 1, n - 1, 3] = 2.5
    return a


SPARSE_SHAPES = [(13, 4, 11, 33, 5, 9, 23, 1), (30, 4, 30, 64, 4, 30, 29, 2), (6, 2, 7, 17, 2, 5, 4, 1),
                 # K = s n above 1024: no tile masks, every step
                 (5, 2, 6, 600, 2, 4, 7, 1)]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", SPARSE_SHAPES)
def test_local_matvec_block_sparse(shape, prec, matvec_mode):
    """Block-sparse operator cores (the stages-2/3 kernel runs only the K steps with a nonzero operator
    tile): against the oracle's dense contraction, and an all-zero core gives exactly zero."""
    L, lib = _lib()
    l1, s, r1, n, S, l2, r2, nv = shape
    _say("local_matvec block-sparse", prec, shape, "mode", matvec_mode)
    dt = torch.float64 if prec == "f64" else torch.float32
    pl, _, pr, x = _rand_local(shape, dt, seed=5)
    for zero in (False, True):
        a = torch.zeros(s, n, n, S, dtype=dt) if zero else torch.tensor(_block_sparse_core(s, n, S, 6), dtype=dt)
        y = torch.full((nv, l1, n, l2), float("nan"), dtype=dt, device=DEV)
        ws = _ws(lib.ttamen_local_matvec_workspace_size(l1, s, r1, n, S, l2, r2, nv) * 2)
        t = [v.to(DEV) for v in (pl, a, pr, x)]
        fn = lib.ttamen_local_matvec if prec == "f64" else lib.ttamen_local_matvec_f32
        code = fn(_p(t[0]), _p(t[1]), _p(t[2]), _p(t[3]), _p(y), l1, s, r1, n, S, l2, r2, nv,
                  _p(ws), ws.numel(), _stream())
        L.check("local_matvec", code)
        torch.cuda.synchronize()
        if zero:
            assert bool((y == 0).all())
            continue
        for v in range(nv):
            want = O.local_product(pr.double().numpy(), pl.double().numpy(), a.double().numpy(),
                                   x[v].double().numpy())
            assert _rel(y[v].cpu().double().numpy(), want) < (1e-13 if prec == "f64" else 2e-5), v


def _bf16(a):
    """Round float32 values to BF16 (round to nearest even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


@pytest.mark.parametrize("shape", SHAPES[:3])
def test_local_matvec_bf16(shape):
    """BF16 products, FP32 accumulation: against a stage-by-stage emulation (operands of each of
    the three contractions rounded to BF16, results stored in FP32) and against FP64."""
    L, lib = _lib()
    l1, s, r1, n, S, l2, r2, nv = shape
    _say("local_matvec bf16", shape)
    pl, a, pr, x = _rand_local(shape, torch.float32, seed=3)
    y = torch.empty(nv, l1, n, l2, dtype=torch.float32, device=DEV)
    ws = _ws(lib.ttamen_local_matvec_workspace_size(l1, s, r1, n, S, l2, r2, nv) * 2)
    t = [v.to(DEV) for v in (pl, a, pr, x)]
    code = lib.ttamen_local_matvec_bf16(_p(t[0]), _p(t[1]), _p(t[2]), _p(t[3]), _p(y), l1, s, r1, n, S, l2, r2,
                                        nv, _p(ws), ws.numel(), _stream())
    L.check("local_matvec_bf16", code)
    torch.cuda.synchronize()
    PL, A4, PR = pl.numpy(), a.numpy(), pr.numpy()
    for v in range(nv):
        X = x[v].numpy()
        t1 = np.einsum("lsr,rnR->lsnR", _bf16(PL).astype(np.float64), _bf16(X).astype(np.float64)).astype(np.float32)
        t2 = np.einsum("smnS,lsnR->lmSR", _bf16(A4).astype(np.float64), _bf16(t1).astype(np.float64)).astype(np.float32)
        emu = np.einsum("lmSR,LSR->lmL", _bf16(t2).astype(np.float64), _bf16(PR).astype(np.float64))
        exact = O.local_product(PR.astype(np.float64), PL.astype(np.float64), A4.astype(np.float64),
                                X.astype(np.float64))
        got = y[v].cpu().double().numpy()
        # FP32 intermediates that differ in the last bit (another accumulation order) can round
        # to neighbouring BF16 values in the next stage: a few 1e-5 of relative difference; a
        # wrong fragment layout or rounding would give O(1)
        assert _rel(got, emu) < 5e-4, v
        assert _rel(got, exact) < 3e-2, v


# ----------------------------------------------------------------- interfaces
@pytest.mark.parametrize("dims", [(6, 2, 5, 16, 3, 7, 4), (84, 4, 84, 64, 4, 84, 84), (1, 1, 1, 9, 2, 3, 2)])
def test_phi_A_updates(dims):
    L, lib = _lib()
    l, s, r, n, S, Lr, R = dims
    _say("phi_A", dims)
    g = torch.Generator().manual_seed(5)
    a = torch.randn(s, n, n, S, dtype=torch.float64, generator=g)
    xl = torch.randn(l, n, Lr, dtype=torch.float64, generator=g)
    xr = torch.randn(r, n, R, dtype=torch.float64, generator=g)
    phi_l = torch.randn(l, s, r, dtype=torch.float64, generator=g)
    phi_r = torch.randn(Lr, S, R, dtype=torch.float64, generator=g)
    ws = _ws(lib.ttamen_phi_workspace_size(l, s, r, n, S, Lr, R))
    out_f = torch.empty(Lr, S, R, dtype=torch.float64, device=DEV)
    out_b = torch.empty(l, s, r, dtype=torch.float64, device=DEV)
    d = [v.to(DEV) for v in (phi_l, xl, a, xr, phi_r)]
    L.check("phi_fwd_A", lib.ttamen_phi_fwd_A(_p(d[0]), _p(d[1]), _p(d[2]), _p(d[3]), _p(out_f), l, s, r, n, S, Lr, R,
                                              _p(ws), ws.numel(), _stream()))
    L.check("phi_bck_A", lib.ttamen_phi_bck_A(_p(d[4]), _p(d[1]), _p(d[2]), _p(d[3]), _p(out_b), l, s, r, n, S, Lr, R,
                                              _p(ws), ws.numel(), _stream()))
    torch.cuda.synchronize()
    want_f = O.phi_fwd_A(phi_l.numpy(), xl.numpy(), a.numpy(), xr.numpy())
    want_b = O.phi_bck_A(phi_r.numpy(), xl.numpy(), a.numpy(), xr.numpy())
    assert _rel(out_f.cpu().numpy(), want_f) < 1e-13
    assert _rel(out_b.cpu().numpy(), want_b) < 1e-13


@pytest.mark.parametrize("dims", [(3, 7, 16, 4, 5), (4, 84, 64, 4, 84), (1, 1, 5, 1, 1)])
def test_phi_rhs_updates(dims):
    L, lib = _lib()
    b, r, n, B, R = dims
    _say("phi_rhs", dims)
    g = torch.Generator().manual_seed(6)
    bc = torch.randn(b, n, B, dtype=torch.float64, generator=g)
    x = torch.randn(r, n, R, dtype=torch.float64, generator=g)
    pl = torch.randn(b, r, dtype=torch.float64, generator=g)
    pr = torch.randn(B, R, dtype=torch.float64, generator=g)
    ws = _ws(64 << 20)
    of = torch.empty(B, R, dtype=torch.float64, device=DEV)
    ob = torch.empty(b, r, dtype=torch.float64, device=DEV)
    d = [v.to(DEV) for v in (pl, bc, x, pr)]
    L.check("phi_fwd_rhs", lib.ttamen_phi_fwd_rhs(_p(d[0]), _p(d[1]), _p(d[2]), _p(of), b, r, n, B, R, _p(ws), ws.numel(), _stream()))
    L.check("phi_bck_rhs", lib.ttamen_phi_bck_rhs(_p(d[3]), _p(d[1]), _p(d[2]), _p(ob), b, r, n, B, R, _p(ws), ws.numel(), _stream()))
    torch.cuda.synchronize()
    assert _rel(of.cpu().numpy(), O.phi_fwd_rhs(pl.numpy(), bc.numpy(), x.numpy())) < 1e-13
    assert _rel(ob.cpu().numpy(), O.phi_bck_rhs(pr.numpy(), bc.numpy(), x.numpy())) < 1e-13


@pytest.mark.parametrize("dims", [(3, 7, 16, 4, 5), (1, 1, 5, 1, 1)])
def test_rhs_assemble(dims):
    """rhs[r,n,R] = scale * sum PL[b,r] bc[b,n,B] PR[B,R] ('br,bnB,BR->rnR', amen_solve.h:303-304)."""
    L, lib = _lib()
    b, r, n, B, R = dims
    _say("rhs_assemble", dims)
    g = torch.Generator().manual_seed(8)
    pl = torch.randn(b, r, dtype=torch.float64, generator=g)
    bc = torch.randn(b, n, B, dtype=torch.float64, generator=g)
    pr = torch.randn(B, R, dtype=torch.float64, generator=g)
    d = [v.to(DEV) for v in (pl, bc, pr)]
    out = torch.empty(r, n, R, dtype=torch.float64, device=DEV)
    ws = _ws(64 << 20)
    L.check("rhs_assemble", lib.ttamen_rhs_assemble(_p(d[0]), _p(d[1]), _p(d[2]), _p(out), b, r, n, B, R, 0.7,
                                                   _p(ws), ws.numel(), _stream()))
    torch.cuda.synchronize()
    ref = 0.7 * np.tensordot(np.tensordot(pl.numpy(), bc.numpy(), axes=([0], [0])), pr.numpy(), axes=([2], [0]))
    assert _rel(out.cpu().numpy(), ref) < 1e-13


def _gmres_problem(seed, r1, s, n, S, r2):
    """A local operator dominated by I (x) (4I - tridiag) (x) I, so that GMRES converges."""
    g = np.random.default_rng(seed)
    PhiL = 0.05 * g.standard_normal((r1, s, r1))
    PhiL[:, 0, :] += np.eye(r1)
    PhiR = 0.05 * g.standard_normal((r2, S, r2))
    PhiR[:, 0, :] += np.eye(r2)
    A = 0.05 * g.standard_normal((s, n, n, S))
    A[0, :, :, 0] += 4 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    rhs = g.standard_normal((r1, n, r2))
    return PhiL, A, PhiR, rhs


@pytest.mark.parametrize("name,m,resets,tol", [("converges", 30, 3, 1e-12), ("capped", 5, 2, 1e-14)])
def test_gmres_solve(name, m, resets, tol):
    """Device GMRES against the oracle's restatement of gmres.h (same Krylov spaces: equal flag,
    iteration count within one per cycle, equal solutions) and, converged, against a dense solve."""
    L, lib = _lib()
    r1, s, n, S, r2 = 6, 2, 8, 2, 5
    N = r1 * n * r2
    PhiL, A, PhiR, rhs = _gmres_problem(4, r1, s, n, S, r2)
    mv = lambda v: O.local_product(PhiR, PhiL, A, v.reshape(r1, n, r2)).reshape(-1)
    xo, flag_o, it_o = O.gmres(mv, rhs.reshape(-1), np.zeros(N), m, tol, resets)
    _say("gmres_solve %s: oracle flag %d iters %d" % (name, flag_o, it_o))
    d = [torch.tensor(v, device=DEV) for v in (PhiL, A, PhiR, rhs)]
    x = torch.empty(N, dtype=torch.float64, device=DEV)
    ws = _ws(lib.ttamen_gmres_solve_workspace_size(r1, s, n, S, r2, m))
    it = ctypes.c_int64(0)
    flag = ctypes.c_int(0)
    L.check("gmres_solve", lib.ttamen_gmres_solve(_p(d[0]), _p(d[1]), _p(d[2]), _p(d[3]), _p(x), r1, s, n, S, r2,
                                                 tol, m, resets, _p(ws), ws.numel(), ctypes.byref(it),
                                                 ctypes.byref(flag), _stream()))
    torch.cuda.synchronize()
    xg = x.cpu().numpy()
    _say("  gpu flag %d iters %d, diff vs oracle %.3e" % (flag.value, it.value, _rel(xg, xo)))
    assert flag.value == flag_o
    assert abs(it.value - it_o) <= resets
    assert _rel(xg, xo) < 1e-10
    if flag_o == 1:
        Bd = np.stack([mv(e) for e in np.eye(N)], axis=1)
        assert _rel(xg, np.linalg.solve(Bd, rhs.reshape(-1))) < 1e-10


@pytest.mark.parametrize("spread", [0.0, 1e-9, 1e-6])
def test_gmres_near_krylov_breakdown(spread):
    """An operator with three clusters of eigenvalues (a Kronecker sum of two small spectra, rotated
    by a random orthogonal matrix; cluster width `spread`): the Krylov space is (nearly) invariant
    after three steps, so the second Gram-Schmidt pass removes almost all of w and the downdated
    norm ||w'||^2 - ||h2||^2 cancels; the close kernel then reduces ||w''|| from w'' itself. The
    solve agrees with the oracle's restatement of gmres.h (explicit norms) and with a dense solve."""
    L, lib = _lib()
    n = 64
    g = np.random.default_rng(17)
    lam = np.add.outer(np.array([0.0, 1.0, 1.0, 0.0]), np.array([1.0, 2.0, 2.0, 1.0])).reshape(-1)  # {1, 2, 3}
    lam = np.tile(lam, n // lam.size) * (1.0 + spread * g.standard_normal(n))
    Q, _ = np.linalg.qr(g.standard_normal((n, n)))
    M = (Q * lam) @ Q.T
    PhiL, PhiR = np.ones((1, 1, 1)), np.ones((1, 1, 1))
    A = M.reshape(1, n, n, 1)
    rhs = g.standard_normal((1, n, 1))
    m, resets, tol = 20, 2, 1e-12
    mv = lambda v: (M @ v.reshape(-1)).reshape(-1)
    xo, flag_o, it_o = O.gmres(mv, rhs.reshape(-1), np.zeros(n), m, tol, resets)
    d = [torch.tensor(v, device=DEV) for v in (PhiL, A, PhiR, rhs)]
    x = torch.empty(n, dtype=torch.float64, device=DEV)
    ws = _ws(lib.ttamen_gmres_solve_workspace_size(1, 1, n, 1, 1, m))
    it, flag = ctypes.c_int64(0), ctypes.c_int(0)
    L.check("gmres_solve", lib.ttamen_gmres_solve(_p(d[0]), _p(d[1]), _p(d[2]), _p(d[3]), _p(x), 1, 1, n, 1, 1,
                                                 tol, m, resets, _p(ws), ws.numel(), ctypes.byref(it),
                                                 ctypes.byref(flag), _stream()))
    torch.cuda.synchronize()
    xg = x.cpu().numpy()
    want = np.linalg.solve(M, rhs.reshape(-1))
    _say("near breakdown (spread %.0e): gpu flag %d iters %d, oracle flag %d iters %d, error vs dense %.2e"
         % (spread, flag.value, it.value, flag_o, it_o, _rel(xg, want)))
    assert np.all(np.isfinite(xg))
    assert flag.value == flag_o == 1
    assert abs(it.value - it_o) <= resets
    assert _rel(xg, want) < 1e-10
    assert _rel(xg, xo) < 1e-10


def test_gemm_autotune():
    """Tuned tiles compute the same products; a solve with the tuner on matches one with it off
    (same ranks, solutions within the solve tolerance)."""
    import torchtt_amd as tt
    L, lib = _lib()
    prev = tt.solvers.gemm_autotune(1)
    try:
        lib.ttamen_gemm_autotune_clear()
        g = torch.Generator().manual_seed(9)
        ws = _ws(256 << 20)
        for (M, N, K, batch) in [(240, 59, 240, 59), (3540, 59, 236, 1), (59, 3540, 59, 4), (56, 56, 5120, 1)]:
            A = torch.randn(batch, M, K, dtype=torch.float64, generator=g).to(DEV)
            B = torch.randn(batch, K, N, dtype=torch.float64, generator=g).to(DEV)
            C = torch.empty(batch, M, N, dtype=torch.float64, device=DEV)
            for _ in range(2):  # tune, then served from the cache
                L.check("gemm_f64", lib.ttamen_gemm_f64(M, N, K, batch, ctypes.c_double(1.0), _p(A), K, 1, M * K,
                                                        _p(B), N, 1, K * N, ctypes.c_double(0.0), _p(C), N, 1,
                                                        M * N, _p(ws), ws.numel(), _stream()))
                torch.cuda.synchronize()
                assert float((C - torch.bmm(A, B)).abs().max() / torch.bmm(A, B).abs().max()) < 1e-13
        st = tt.solvers.gemm_autotune_stats()
        _say("autotune stats", st)
        assert st["buckets_tuned"] == 4 and st["launches_from_cache"] >= 4
        from torchtt_amd import workloads as W
        A, b = W.laplace_potential(5, 16, seed=1)
        Ag, bg = tt.TT([c.to(DEV) for c in A]), tt.TT([c.to(DEV) for c in b])
        torch.manual_seed(0)
        x1 = tt.solvers.amen_solve(Ag, bg, eps=1e-10, rmax=30, verbose=False)
        tt.solvers.gemm_autotune(0)
        torch.manual_seed(0)
        x0 = tt.solvers.amen_solve(Ag, bg, eps=1e-10, rmax=30, verbose=False)
        _say("autotune solve: ranks %s / %s diff %.3e" % (list(x1.R), list(x0.R), float((x1 - x0).norm() / x0.norm())))
        assert list(x1.R) == list(x0.R)
        assert float((x1 - x0).norm() / x0.norm()) < 1e-8
    finally:
        tt.solvers.gemm_autotune(prev)
        lib.ttamen_gemm_autotune_clear()


def test_device_info():
    L, lib = _lib()
    buf = ctypes.create_string_buffer(256)
    L.check("device_info", lib.ttamen_device_info(buf, 256))
    txt = buf.value.decode()
    _say("device_info:", txt, "|", lib.ttamen_version().decode())
    assert "CUs" in txt and "gfx" in txt


# ----------------------------------------------------------------- QR / SVD
def _qr_case(kind, m, n, seed):
    return qr_case(kind, m, n, seed)


@pytest.mark.parametrize("kind,m,n", [("rand", 200, 30), ("rand", 5376, 84), ("rand", 30, 30), ("rand", 10, 25),
                                      ("zero_cols", 100, 12), ("dependent", 80, 10), ("graded", 300, 40),
                                      ("tiny_tail", 1000, 40), ("rand", 1, 1), ("rand", 256, 4),
                                      ("rand", 8448, 136),
                                      # TRSM column blocks: the second wave idle (6, 8), a ragged block (17)
                                      ("rand", 100, 6), ("graded", 300, 8), ("dependent", 500, 17),
                                      # single-workgroup path (up to 4 columns)
                                      ("zero_cols", 300, 4), ("dependent", 64, 3), ("graded", 256, 4),
                                      ("rand", 4352, 4), ("rand", 7, 2), ("zero_cols", 40, 2),
                                      # numerically rank-deficient and near-square inputs
                                      ("rank_deficient", 16, 12), ("rank_deficient", 120, 110),
                                      ("rank_deficient", 3000, 60), ("rank_deficient", 64, 4),
                                      ("square_zero_col", 148, 137), ("square_zero_col", 5, 4),
                                      # more than 140 columns: Householder QR in device memory
                                      ("rand", 3000, 150), ("rank_deficient", 400, 200), ("zero_cols", 1000, 160),
                                      ("square_zero_col", 180, 170), ("graded", 600, 141)])
def test_qr(kind, m, n):
    L, lib = _lib()
    _say("qr", kind, m, n)
    A = _qr_case(kind, m, n, m + n)
    k = min(m, n)
    Ad = torch.from_numpy(A).to(DEV)
    Q = torch.empty(m, k, dtype=torch.float64, device=DEV)
    R = torch.empty(k, n, dtype=torch.float64, device=DEV)
    ws = _ws(lib.ttamen_dense_workspace_size(m, n))
    L.check("qr", lib.ttamen_qr(_p(Ad), m, n, _p(Q), _p(R), _p(ws), ws.numel(), _stream()))
    Qh, Rh = Q.cpu().numpy(), R.cpu().numpy()
    assert np.linalg.norm(Qh.T @ Qh - np.eye(k)) < 1e-13 * max(1, np.sqrt(k)) * 10
    assert np.linalg.norm(Qh @ Rh - A) <= 1e-14 * np.sqrt(m * n) * 10 * np.linalg.norm(A) + 1e-300


@pytest.mark.parametrize("case", ["rand", "in_span", "zero_and_duplicate", "nearly_in_span"])
@pytest.mark.parametrize("m,r,p", [(3328, 52, 4), (300, 40, 3), (44, 40, 4), (64, 1, 1)])
def test_qr_extend(case, m, r, p):
    """QR of [U | Z] with U orthonormal (the solver's enrichment QR): Q keeps U as its first columns, is
    orthonormal, and Q R = A; Z columns in span(U), zero or repeated are completed against all of Q."""
    L, lib = _lib()
    _say("qr_extend", case, m, r, p)
    g = np.random.default_rng(m + r + p)
    U, _ = np.linalg.qr(g.standard_normal((m, r)))
    Z = g.standard_normal((m, p))
    if case == "in_span":
        Z[:, 0] = U @ g.standard_normal(r)
    elif case == "zero_and_duplicate":
        Z[:, 0] = 0.0
        if p > 1:
            Z[:, p - 1] = Z[:, 0 if p == 2 else 1]
    elif case == "nearly_in_span":
        Z[:, 0] = U @ g.standard_normal(r) + 1e-12 * g.standard_normal(m)
    A = np.ascontiguousarray(np.concatenate([U, Z], axis=1))
    n = r + p
    Ad = torch.from_numpy(A).to(DEV)
    Q = torch.empty(m, n, dtype=torch.float64, device=DEV)
    R = torch.empty(n, n, dtype=torch.float64, device=DEV)
    ws = _ws(lib.ttamen_dense_workspace_size(m, n))
    L.check("qr_extend", lib.ttamen_qr_extend(_p(Ad), m, r, p, _p(Q), _p(R), _p(ws), ws.numel(), _stream()))
    Qh, Rh = Q.cpu().numpy(), R.cpu().numpy()
    assert np.array_equal(Qh[:, :r], U)
    assert np.linalg.norm(Qh.T @ Qh - np.eye(n)) < 1e-13 * np.sqrt(n) * 10
    assert np.linalg.norm(Qh @ Rh - A) <= 1e-14 * np.sqrt(m * n) * 10 * np.linalg.norm(A)


@pytest.mark.parametrize("pert", [0.0, 1e-3, 1e-2])
@pytest.mark.parametrize("m,n", [(3328, 52), (600, 80)])
def test_qr_nearly_orthonormal_input(m, n, pert):
    """Inputs already orthonormal up to `pert`: a Cholesky-QR pass whose Gram G has n max|G - I| < 0.1 is
    unshifted and the last (0 and 1e-3 here at n = 52; 1e-2 and the n = 80 cases with 1e-3 take the shifted
    passes first). Q must come out orthonormal to working precision and Q R = A either way; for an orthonormal
    input R is the identity up to rounding."""
    L, lib = _lib()
    _say("qr nearly orthonormal", m, n, pert)
    rng = np.random.default_rng(m + n)
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    A = U + pert * rng.standard_normal((m, n)) / np.sqrt(m)
    Ad = torch.from_numpy(A).to(DEV)
    Q = torch.empty(m, n, dtype=torch.float64, device=DEV)
    R = torch.empty(n, n, dtype=torch.float64, device=DEV)
    ws = _ws(lib.ttamen_dense_workspace_size(m, n))
    L.check("qr", lib.ttamen_qr(_p(Ad), m, n, _p(Q), _p(R), _p(ws), ws.numel(), _stream()))
    Qh, Rh = Q.cpu().numpy(), R.cpu().numpy()
    assert np.linalg.norm(Qh.T @ Qh - np.eye(n)) < 1e-13 * np.sqrt(n)
    assert np.linalg.norm(Qh @ Rh - A) <= 1e-14 * np.sqrt(m * n) * 10 * np.linalg.norm(A)
    if pert == 0.0:
        assert np.max(np.abs(np.abs(Rh) - np.eye(n))) < 1e-13


def test_qr_extend_refuses_shapes():
    L, lib = _lib()
    m, r, p = 40, 38, 3  # r + p > m
    A = torch.zeros(m, r + p, dtype=torch.float64, device=DEV)
    Q = torch.empty(m, r + p, dtype=torch.float64, device=DEV)
    R = torch.empty(r + p, r + p, dtype=torch.float64, device=DEV)
    ws = _ws(lib.ttamen_dense_workspace_size(m, r + p))
    assert lib.ttamen_qr_extend(_p(A), m, r, p, _p(Q), _p(R), _p(ws), ws.numel(), _stream()) == 1


@pytest.mark.parametrize("kind,m,n", [("rand", 200, 30), ("rand", 5376, 84), ("rand", 4, 256), ("rand", 30, 30),
                                      ("graded", 300, 40), ("zero_cols", 60, 9), ("rand", 1, 7), ("rand", 9, 1),
                                      ("rand", 1000, 136), ("rand", 4, 256), ("rand", 256, 4), ("zero_cols", 128, 4),
                                      ("rank_deficient", 16, 12), ("rank_deficient", 12, 16),
                                      ("rank_deficient", 192, 14), ("square_zero_col", 148, 137),
                                      # rank above 140: one-sided Jacobi in device memory (zero columns: the
                                      # completion of zero singular values)
                                      ("rand", 2000, 150), ("rank_deficient", 300, 180), ("rand", 160, 400),
                                      ("zero_cols", 500, 150), ("square_zero_col", 170, 170)])
def test_svd(kind, m, n):
    L, lib = _lib()
    _say("svd", kind, m, n)
    A = _qr_case(kind, m, n, 3 * m + n)
    k = min(m, n)
    Ad = torch.from_numpy(A).to(DEV)
    U = torch.empty(m, k, dtype=torch.float64, device=DEV)
    s = torch.empty(k, dtype=torch.float64, device=DEV)
    V = torch.empty(n, k, dtype=torch.float64, device=DEV)
    ws = _ws(lib.ttamen_dense_workspace_size(m, n))
    L.check("svd", lib.ttamen_svd(_p(Ad), m, n, _p(U), _p(s), _p(V), _p(ws), ws.numel(), _stream()))
    Uh, sh, Vh = U.cpu().numpy(), s.cpu().numpy(), V.cpu().numpy()
    s_np = np.linalg.svd(A, compute_uv=False)
    smax = max(s_np[0], 1e-300)
    assert np.all(np.diff(sh) <= 1e-15 * smax)
    assert np.max(np.abs(sh - s_np)) <= 1e-13 * smax * max(1, np.sqrt(k))
    # Jacobi stops once every pair of columns has |cosine| <= k u, so ||U^T U - I||_F is bounded by
    # about k^2 u: 1e-12 up to k = 134, k^2 2^-52 above (7.2e-12 at k = 180)
    orth = max(1e-12, k * k * 2.0 ** -52)
    assert np.linalg.norm(Uh.T @ Uh - np.eye(k)) < orth
    assert np.linalg.norm(Vh.T @ Vh - np.eye(k)) < orth
    assert np.linalg.norm((Uh * sh) @ Vh.T - A) <= 1e-13 * np.sqrt(m * n) * np.linalg.norm(A)


@pytest.mark.parametrize("n", [1, 7, 31, 32, 33, 120, 499])
def test_dense_solve(n):
    L, lib = _lib()
    _say("dense_solve", n)
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n)) + n * 0.05 * np.eye(n)
    rhs = rng.standard_normal(n)
    Bd = torch.from_numpy(B.copy()).to(DEV)
    rd = torch.from_numpy(rhs.copy()).to(DEV)
    xd = torch.empty(n, dtype=torch.float64, device=DEV)
    ws = _ws(lib.ttamen_dense_solve_workspace_size(n))
    L.check("dense_solve", lib.ttamen_dense_solve(_p(Bd), _p(rd), _p(xd), n, _p(ws), ws.numel(), _stream()))
    torch.cuda.synchronize()
    x = xd.cpu().numpy()
    want = np.linalg.solve(B, rhs)
    assert _rel(x, want) < 1e-11 * max(1.0, np.linalg.cond(B) / 1e3)
