import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def pkg():
    from __graft_entry__ import load_package
    return load_package()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.lib()
    return O


def _arr(entry):
    return np.array([float.fromhex(h) for h in entry["hex"]]).reshape(entry["shape"], order="F")


@pytest.fixture(scope="session")
def reftests():
    with open(os.path.join(GOLDEN, "reftests.json")) as f:
        raw = json.load(f)
    return {k: {"X": _arr(v["X"]), "H": _arr(v["H"]), "sha256": v["sha256"]} for k, v in raw.items()}


@pytest.fixture(scope="session")
def iteration_tables():
    with open(os.path.join(GOLDEN, "iterations.json")) as f:
        return json.load(f)


# fixture name -> (script, n): test/test_1D.jl (NX=64), test/test_2D.jl (NX=NY=64), test/shared.jl (32^3)
FIXTURE_SCRIPTS = {
    "test_diff_1D_lin": ("diff_1D_lin", 64),
    "test_diff_1D_linstep": ("diff_1D_linstep", 64),
    "test_diff_1D_nonlin": ("diff_1D_nonlin", 64),
    "test_diff_1D_react": ("diff_1D_react", 64),
    "test_diff_2D_lin": ("diff_2D_lin", 64),
    "test_diff_2D_linstep": ("diff_2D_linstep", 64),
    "test_diff_2D_nonlin": ("diff_2D_nonlin", 64),
    "test_diff_2D_nonlin_perf": ("diff_2D_nonlin_perf", 64),
    "test_diff_2D_react": ("diff_2D_react", 64),
    "test_diff_3D_lin_multixpu": ("diff_3D_lin_multixpu", 32),
    "test_diff_3D_linstep_multixpu": ("diff_3D_linstep_multixpu", 32),
    "test_diff_3D_nonlin_multixpu": ("diff_3D_nonlin_multixpu", 32),
    "test_diff_3D_nonlin_multixpu_perf": ("diff_3D_nonlin_multixpu_perf", 32),
    "test_diff_3D_nonlin_perf": ("diff_3D_nonlin_perf", 32),
}


def sample(H, n, ndim):
    """H[inds, inds(, inds)] with inds = Int.(ceil.(LinRange(1, n, 12))) (test/test_1D.jl:23)."""
    t = np.arange(12) / 11.0
    r = (1.0 - t) * 1.0 + t * float(n)
    inds = np.ceil(r).astype(int) - 1
    return H[np.ix_(*([inds] * ndim))], inds


def isapprox_norm(a, b):
    """Julia isapprox(a, b) with default rtol = sqrt(eps): norm(a-b) <= rtol*max(norm(a), norm(b))."""
    return np.linalg.norm(a - b) <= np.sqrt(np.finfo(float).eps) * max(np.linalg.norm(a), np.linalg.norm(b))


# The oracle's per-step PT iteration counts of the 1D scripts at n = 16384 with its own initial condition
# (glibc exp) and row-ordered residual sums; recomputed on the CPU by
# tests/test_oracle_golden.py::test_oracle_1d_16384_counts. At this size the residual sits on a roundoff
# plateau within a few percent of tol for hundreds of iterations, so the stopping check depends on the
# last-ulp history: the reference's published totals (63530, 64460, 30430; output/out_diff_1D_*.txt:10)
# come from other arithmetic (Julia's exp, CUDA, CUBLAS nrm2) and are not reproducible at the count level.
ORACLE_1D_16384 = {
    "1D_lin": [13790, 12910, 12590, 12300, 12080],
    "1D_linstep": [14970, 12710, 12340, 12270, 12190],
    "1D_nonlin": [7410, 6290, 5830, 5590, 5340],
}
