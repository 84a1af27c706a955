"""Shared test fixtures. Tests marked `gpu` need a ROCm GPU and the built libttamen.so;
everything else runs on the CPU (oracle against golden vectors, host logic, ABI exports)."""
import glob
import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and libttamen.so")


def load_golden(name):
    """Golden case written by tests/golden/make_golden.py from the reference's C++ solver."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    d = int(z["d"])
    case = {k: z[k] for k in z.files if not k[0] in "Abzx" or k in ("d",)}
    case["A"] = [z["A%d" % k] for k in range(d)]
    case["b"] = [z["b%d" % k] for k in range(d)]
    case["z"] = [z["z%d" % k] for k in range(d)]
    case["x"] = [z["x%d" % k] for k in range(d)]
    case["x0"] = [z["x0_%d" % k] for k in range(d)] if ("x0_0" in z.files) else None
    for key in ("d", "seed", "nswp", "rmax", "kickrank", "kick2"):
        case[key] = int(z[key])
    case["eps"] = float(z["eps"])
    return case


GOLDEN_CASES = ["laplace_d6_n16_eps1e-8", "laplace_d6_n16_eps1e-10", "random_456", "potential_d4_n16",
                "laplace_d5_n12_x0"]

# Sweeps the oracle takes on each fixture (the last, non-enriching sweep included). The CPU
# suite checks the oracle against these and the GPU suite checks the native solver: a solver
# whose local linear algebra loses accuracy needs more sweeps on the symmetric Laplace cases
# before its residual proxy falls below eps, even when its final ranks still agree.
GOLDEN_SWEEPS = {"laplace_d6_n16_eps1e-8": 5, "laplace_d6_n16_eps1e-10": 6, "random_456": 5,
                 "potential_d4_n16": 11, "laplace_d5_n12_x0": 5}


def load_ref():
    """The reference's own C++ AMEn (oracle/_ref, built from its sources), or None."""
    so = glob.glob(os.path.join(ROOT, "oracle", "_ref", "torchttcpp*.so"))
    if not so:
        return None
    spec = importlib.util.spec_from_file_location("torchttcpp", so[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
