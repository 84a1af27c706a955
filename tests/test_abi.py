"""The C-ABI library loads and exports every entry point include/ttamen.h declares; the
host-side size queries work without a GPU."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ttamen.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ttamen_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    from torchtt_amd import _lib
    assert _declared() == sorted(_lib.exported_names())


def test_library_exports_every_declared_symbol():
    from torchtt_amd import _lib
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert b"gfx950" in lib.ttamen_version()


def test_workspace_queries_without_gpu():
    from torchtt_amd import _lib
    lib = _lib.load()
    d = 10
    N = _lib.i64_array([64] * d)
    rA = _lib.i64_array([1] + [4] * (d - 1) + [1])
    rb = _lib.i64_array([1] + [4] * (d - 1) + [1])
    ws = lib.ttamen_amen_solve_workspace_size(d, N, rA, rb, None, 80, 4, 0, 40, 500, 0)
    assert ws > 0
    # solution buffer: cores at their rank bounds min(r_{k} n, rmax + kick, ...)
    out = lib.ttamen_amen_solve_output_elems(d, N, None, 80, 4, 0)
    R = [1]
    for k in range(d):
        right = 64 ** (d - k - 1)
        R.append(1 if k == d - 1 else min(R[-1] * 64, 84, right + 4))
    assert out == sum(R[k] * 64 * R[k + 1] for k in range(d))
    # invalid input: size 0 and a message
    bad = _lib.i64_array([1] + [4] * (d - 1) + [2])  # boundary rank must be 1
    assert lib.ttamen_amen_solve_workspace_size(d, N, bad, rb, None, 80, 4, 0, 40, 500, 0) == 0
    assert b"boundary" in lib.ttamen_last_error()


def test_complex_cores_are_refused_before_any_device_work():
    """The native sweeps compute in real FP64; a complex TT is refused with InvalidArguments (a cast
    would silently drop the imaginary parts). Runs without a GPU: the check precedes device work."""
    import pytest
    import torch
    import torchtt_amd as tt
    from torchtt_amd.errors import InvalidArguments
    N = [3, 4, 3]
    A = tt.TT([torch.eye(n, dtype=torch.complex128).reshape(1, n, n, 1) for n in N])
    Ar = tt.TT([torch.eye(n, dtype=torch.float64).reshape(1, n, n, 1) for n in N])
    b = tt.TT([torch.ones(1, n, 1, dtype=torch.complex128) for n in N])
    br = tt.TT([torch.ones(1, n, 1, dtype=torch.float64) for n in N])
    with pytest.raises(InvalidArguments):
        tt.solvers.amen_solve(A, br)
    with pytest.raises(InvalidArguments):
        tt.solvers.amen_solve(Ar, b)
    with pytest.raises(InvalidArguments):
        tt.solvers.amen_solve(Ar, br, x0=b)
    with pytest.raises(InvalidArguments):
        tt.solvers.amen_divide(b, br)
    with pytest.raises(InvalidArguments):
        tt.amen_mv(A, br)
    with pytest.raises(InvalidArguments):
        tt.solvers.round_tt(b)
    with pytest.raises(InvalidArguments):
        tt.solvers.tt_full(b)
    from torchtt_amd._dmrg import dmrg_matvec
    with pytest.raises(InvalidArguments):
        dmrg_matvec(A, br)
