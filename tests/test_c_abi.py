"""The C ABI from a plain C caller (tests/c/abi_solve.c, built by tests/c/Makefile): include/ptd.h compiles
as strict C11 (gcc -std=c11 -pedantic -Werror) and the caller links against libptd.so (CPU); on the GPU its
ptd_solve run gives the iteration counts, errors and final H of the Python binding's run of the same case,
bit for bit. The C caller stands for any C-level host, such as a Julia `ccall` binding."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

C_DIR = os.path.join(ROOT, "tests", "c")
BUILT = os.path.join(C_DIR, "libabi_solve.so")
LIB_DIR = os.path.join(ROOT, "pseudotransientdiffusion.jl_amd")


def test_header_compiles_as_c11_and_the_caller_links(tmp_path, pkg):
    pkg.load()  # libptd.so built and loadable
    out = tmp_path / "libabi_solve.so"
    r = subprocess.run(["make", "-B", "-C", C_DIR, f"OUT={out}", f"RPATH={LIB_DIR}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    lib = ctypes.CDLL(str(out))
    assert hasattr(lib, "abi_solve")


def _abi_solve(lib, params, H0, nt):
    f64p = ctypes.POINTER(ctypes.c_double)
    lib.abi_solve.restype = ctypes.c_int32
    lib.abi_solve.argtypes = [ctypes.c_void_p, f64p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), f64p, f64p]
    iters = np.zeros(nt, dtype=np.int64)
    errs = np.zeros(nt, dtype=np.float64)
    H = np.zeros_like(H0, order="F")
    st = lib.abi_solve(ctypes.addressof(params), H0.ctypes.data_as(f64p), nt,
                       iters.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), errs.ctypes.data_as(f64p),
                       H.ctypes.data_as(f64p))
    return st, iters, errs, H


@pytest.mark.gpu
@pytest.mark.parametrize("script,shape", [("diff_3D_nonlin_multixpu", (45, 38, 29)), ("diff_2D_lin", (64, 60)),
                                          ("diff_1D_nonlin", (128,))])
def test_c_caller_solves_like_the_python_binding(script, shape, pkg):
    if not os.path.exists(BUILT):
        pytest.skip("tests/c/libabi_solve.so is not built (python -c 'import __graft_entry__ as g; g.build()')")
    print(f"[gpu] C caller {script} {shape}", flush=True)
    lib = ctypes.CDLL(BUILT)
    spec = pkg.SCRIPTS[script]
    params = pkg.make_params(spec, *shape)
    H0 = np.asfortranarray(pkg.initial_condition(spec, *shape), dtype=np.float64)
    nt = 3
    st, iters, errs, H = _abi_solve(lib, params, H0, nt)
    assert st == 0
    with pkg.Context(params) as ctx:
        ctx.set("H", H0)
        ctx.set("Hold", H0)
        it_py, err_py = ctx.solve(nt)
        H_py = ctx.get("H")
    assert list(iters) == [int(v) for v in it_py]
    assert np.array_equal(errs, np.asarray(err_py, dtype=np.float64))
    assert np.array_equal(H, H_py)
    assert np.all(errs <= params.tol)
