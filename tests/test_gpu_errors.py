"""Allocation failures on the device: the context, gradient and batch entry points report
PTD_ERR_NOMEM when their buffers exceed the device memory, and the calls that follow (new contexts,
kernel launches, the same context) run as if the failed call had not happened. Each request is far
beyond the 288 GB of HBM, so hipMalloc refuses it at once."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_context_allocation_failure_reports_nomem(pkg, orc):
    print("[gpu] context of 4096^3 float64 (550 GB per field)", flush=True)
    p = pkg.make_params(pkg.SCRIPTS["diff_3D_lin_multixpu"], 4096, 4096, 4096)
    with pytest.raises(pkg.PTDError) as ei:
        pkg.Context(p)
    assert ei.value.status == pkg._abi.ERR_NOMEM
    # a later solve launches its kernels without inheriting the failed allocation's error
    res = pkg.run_script("diff_3D_lin_multixpu", 16, 16, 16)
    ref = orc.run_script("diff_3D_lin_multixpu", 16, 16, 16)
    assert np.array_equal(np.asarray(res.iters), np.asarray(ref["iters"]))
    assert np.max(np.abs(res.H - ref["H"])) <= 1e-12 * np.max(np.abs(ref["H"]))


def _step_ctx(pkg, n):
    spec = pkg.SCRIPTS["diff_3D_linstep_multixpu"]
    p = pkg.make_params(spec, n, n, n, arith=1)
    ctx = pkg.Context(p)
    H0 = pkg.initial_condition(spec, n, n, n)
    ctx.set("H", H0)
    ctx.set("Hold", H0)
    ctx.set("D", pkg.step_field(spec, n, n, n))
    return ctx, H0


def test_gradient_allocation_failure_reports_nomem_and_keeps_the_context(pkg):
    n, nt_huge = 48, 4_000_000  # checkpoints: 4e6 x 48^3 x 8 B = 3.5 TB
    print(f"[gpu] gradient checkpoints for nt={nt_huge} at {n}^3", flush=True)
    ctx, H0 = _step_ctx(pkg, n)
    hobs = np.asfortranarray(H0 * 1.01)
    with ctx:
        lib = pkg.load()
        st = lib.ptd_grad_solve(ctx.handle, nt_huge, hobs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                hobs.size, None, None, None, None, None)
        assert st == pkg._abi.ERR_NOMEM, lib.ptd_last_error()
        after = ctx.grad_solve(2, hobs)
    # the failed call changed nothing: the same gradient as from a fresh context
    fresh, _ = _step_ctx(pkg, n)
    with fresh:
        ref = fresh.grad_solve(2, hobs)
    assert after["loss"] == ref["loss"]
    assert np.array_equal(after["grad_D"], ref["grad_D"])
    assert np.array_equal(after["iters_fwd"], ref["iters_fwd"])


def test_batch_allocation_failure_reports_nomem(pkg):
    big = pkg.make_params(pkg.SCRIPTS["diff_2D_lin"], 200_000, 200_000)  # 320 GB per field
    small = pkg.make_params(pkg.SCRIPTS["diff_2D_lin"], 64, 64)
    print("[gpu] batch with a 200000^2 member", flush=True)
    with pytest.raises(pkg.PTDError) as ei:
        pkg.Batch([small, big])
    assert ei.value.status == pkg._abi.ERR_NOMEM
    with pkg.Batch([small]) as b:
        b.set(0, "H", pkg.initial_condition(pkg.SCRIPTS["diff_2D_lin"], 64, 64))
        b.set(0, "Hold", pkg.initial_condition(pkg.SCRIPTS["diff_2D_lin"], 64, 64))
        iters, errs, lat, status, total = b.solve(1)
    assert status[0] == 0 and iters[0, 0] > 0 and errs[0, 0] <= small.tol
