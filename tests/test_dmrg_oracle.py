"""The oracle's DMRG matvec (restatement of cpp/dmrg_mv.h) and its normal stream, pinned by
published SplitMix64 outputs, dense products, the reference's own test structure
(tests/test_algebra_2.py:39-82), identity and zero operands, and the rank rule of the last sweep
(truncation rank of the exact product plus the kick)."""
import numpy as np
import pytest

import amen_oracle as O


def _cores(rng, ranks, modes):
    return [rng.standard_normal((ranks[k],) + tuple(modes[k]) + (ranks[k + 1],)) for k in range(len(modes))]


def _add(x, y):
    """Cores of x + y for vector or operator cores (block concatenation)."""
    shp = [c.shape for c in x]
    x3 = [c.reshape(c.shape[0], -1, c.shape[-1]) for c in x]
    y3 = [c.reshape(c.shape[0], -1, c.shape[-1]) for c in y]
    s = O.tt_add(x3, y3)
    return [c.reshape((c.shape[0],) + tuple(sh[1:-1]) + (c.shape[-1],)) for c, sh in zip(s, shp)]


def _full_op(A):
    t = A[0]
    for a in A[1:]:
        t = np.tensordot(t, a, axes=([t.ndim - 1], [0]))
    t = t.reshape(t.shape[1:-1])
    d = t.ndim // 2
    t = t.transpose(list(range(0, 2 * d, 2)) + list(range(1, 2 * d, 2)))
    return t.reshape(int(np.prod(t.shape[:d])), -1)


def _guess(rng, M):
    d = len(M)
    return _cores(rng, [1] + (d - 1) * [2] + [1], [(m,) for m in M])


def test_splitmix64_known_answers():
    """SplitMix64 from state 1234567: the first five outputs of the published generator."""
    g = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        states = np.uint64(1234567) + np.arange(5, dtype=np.uint64) * g
    out = [int(v) for v in O._splitmix64(states)]
    assert out == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                   4593380528125082431, 16408922859458223821]


def test_randn_stream_moments():
    z = O.randn_stream(11, 5, 400000)
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1.0) < 0.01
    assert abs(np.mean(np.abs(z) < 1.0) - 0.6826894921370859) < 0.005  # P(|Z| < 1)
    w = O.randn_stream(11, 6, 400000)
    assert abs(np.mean(z * w)) < 0.01  # different streams: uncorrelated
    assert np.array_equal(O.randn_stream(11, 5, 1000), z[:1000])  # counter-based: prefix-stable


@pytest.mark.parametrize("nm", [0, 2])
def test_dmrg_dense_small(nm):
    """y against the dense product (square and rectangular operator cores)."""
    rng = np.random.default_rng(3 + nm)
    d, n = 4, 3
    A = _cores(rng, [1, 2, 3, 2, 1], [(n + nm, n)] * d)
    x = _cores(rng, [1, 3, 2, 3, 1], [(n,)] * d)
    tr = {}
    y = O.dmrg_mv(A, x, _guess(rng, [n + nm] * d), seed=5, trace=tr)
    ex = _full_op(A) @ O.tt_full(x).reshape(-1)
    err = np.linalg.norm(O.tt_full(y).reshape(-1) - ex) / np.linalg.norm(ex)
    assert err < 1e-13
    assert tr["sweeps"] < 20


def test_dmrg_reference_structure():
    """tests/test_algebra_2.py:39-60 in FP64 at n = 6, d = 6: (A + A) fast_matvec (4 x) against 8 A x."""
    rng = np.random.default_rng(7)
    d, n = 6, 6
    A = _cores(rng, [1] + (d - 1) * [3] + [1], [(n, n)] * d)
    x = _cores(rng, [1] + (d - 1) * [5] + [1], [(n,)] * d)
    Am = _add(A, A)
    xm = _add(x, x)
    xm = _add(xm, xm)
    y = O.dmrg_mv(Am, xm, _guess(rng, [n] * d), seed=1)
    ref = O.tt_matvec(A, x)
    ref[0] = 8.0 * ref[0]
    err = O.tt_norm(O.tt_add(y, ref, 1.0, -1.0)) / O.tt_norm(ref)
    assert err < 1e-12


def test_dmrg_final_ranks():
    """The last sweep truncates y_k y_k+1 to the rank of the product at eps and adds the kick:
    ry_k+1 = min(rank_k+1 + kickrank, ry_k M_k), with rank_k the unfolding ranks of A x."""
    rng = np.random.default_rng(9)
    d, n, kick = 5, 3, 2
    A = _cores(rng, [1, 2, 2, 2, 2, 1], [(n, n)] * d)
    x = _cores(rng, [1, 2, 1, 2, 2, 1], [(n,)] * d)
    y = O.dmrg_mv(A, x, _guess(rng, [n] * d), kickrank=kick, seed=2)
    full = (_full_op(A) @ O.tt_full(x).reshape(-1)).reshape([n] * d)
    ry = [1]
    for k in range(1, d):
        rk = np.linalg.matrix_rank(full.reshape(n ** k, -1), tol=1e-10 * np.linalg.norm(full))
        ry.append(min(rk + kick, ry[-1] * n))
    assert [c.shape[0] for c in y] == ry


def test_dmrg_identity_and_zero():
    rng = np.random.default_rng(13)
    d, n = 4, 4
    I = [np.eye(n).reshape(1, n, n, 1) for _ in range(d)]
    x = _cores(rng, [1, 3, 4, 2, 1], [(n,)] * d)
    y = O.dmrg_mv(I, x, _guess(rng, [n] * d), seed=3)
    assert O.tt_norm(O.tt_add(y, x, 1.0, -1.0)) / O.tt_norm(x) < 1e-13
    z = [np.zeros_like(c) for c in x]
    y = O.dmrg_mv(I, z, _guess(rng, [n] * d), nswp=4, seed=3)
    assert O.tt_norm(y) == 0.0


def test_dmrg_rmax_caps_ranks():
    rng = np.random.default_rng(17)
    d, n = 5, 4
    A = _cores(rng, [1, 3, 3, 3, 3, 1], [(n, n)] * d)
    x = _cores(rng, [1, 4, 4, 4, 4, 1], [(n,)] * d)
    y = O.dmrg_mv(A, x, _guess(rng, [n] * d), rmax=5, kickrank=2, nswp=6, seed=4)
    assert max(c.shape[0] for c in y) <= 5 + 2
