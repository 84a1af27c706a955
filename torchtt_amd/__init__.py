"""torchtt_amd -- MI355X-native AMEn linear solver for tensor-train systems.

Drop-in for `torchtt.solvers.amen_solve(A, b, ...)` (ion-g-ion/torchTT): the same call
and the same TT return type, with the solve running in hand-written HIP kernels for
gfx950 (libttamen.so, C ABI in include/ttamen.h).
"""
from ._tt import TT
from ._extras import eye, ones, zeros, random, randn, kron, dot, meshgrid, elementwise_divide
from ._amen import amen_mv, amen_mm
from . import errors
from . import solvers
from . import workloads
from . import serving

__all__ = ["TT", "eye", "ones", "zeros", "random", "randn", "kron", "dot", "meshgrid", "elementwise_divide",
           "amen_mv", "amen_mm", "errors", "solvers", "workloads", "serving"]
