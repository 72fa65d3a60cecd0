"""Error paths on the device.

Allocation failures: the context, gradient and batch entry points report PTD_ERR_NOMEM when their
buffers exceed the device memory, and the calls that follow (new contexts, kernel launches, the same
context) run as if the failed call had not happened. Each request is far beyond the 288 GB of HBM, so
hipMalloc refuses it at once.

NaN: a NaN in the initial condition reaches the residual at the first check, and the solve stops there
with PTD_ERR_NAN, as the scripts stop with error("NaN") (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:118),
with the same iteration counts, errors and fields as the oracle. NaN values are ordinary data to the
kernels: nothing faults."""
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


NAN_CASES = [("diff_2D_nonlin", (40, 36)), ("diff_3D_linstep_multixpu", (24, 20, 18)),
             ("diff_3D_nonlin_multixpu", (26, 22, 20))]


def _nan_ic(orc, script, shape):
    nx, ny, nz = (list(shape) + [1, 1])[:3]
    ref = orc.run_script(script, nx, ny, nz)
    H0 = np.array(ref["H0"], order="F")
    H0[tuple(s // 2 for s in shape)] = np.nan
    return orc.make_case(orc.SCRIPTS[script], nx, ny, nz), H0, ref["D"]


@pytest.mark.parametrize("script,shape", NAN_CASES)
def test_nan_stops_the_solve_like_the_oracle(script, shape, pkg, orc):
    print(f"[gpu] NaN initial condition {script} {shape}", flush=True)
    case, H0, D = _nan_ic(orc, script, shape)
    Hr, Hor = H0.copy(order="F"), H0.copy(order="F")
    it_r, er_r, st_r = orc.solve(case, Hr, Hor, D)
    assert st_r == 1 and it_r[0] > 0
    nt = len(it_r)
    p = pkg.make_params(pkg.SCRIPTS[script], *shape)  # EXACT arithmetic: the oracle's bits
    lib = pkg.load()
    with pkg.Context(p) as ctx:
        ctx.set("H", H0)
        ctx.set("Hold", H0)
        if D is not None:
            ctx.set("D", D)
        iters = np.zeros(nt, dtype=np.int64)
        errs = np.zeros(nt)
        st = lib.ptd_solve(ctx.handle, nt, iters.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                           errs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        assert st == pkg._abi.ERR_NAN, lib.ptd_last_error()
        assert np.array_equal(iters, it_r)
        assert np.isnan(errs[0]) and np.isnan(er_r[0])
        H = ctx.get("H")
        Hold = ctx.get("Hold")
    assert np.array_equal(np.isnan(H), np.isnan(Hr))
    assert np.array_equal(H, Hr, equal_nan=True)
    assert np.array_equal(Hold, Hor, equal_nan=True)


def test_nan_member_leaves_the_batch_intact(pkg, orc):
    print("[gpu] batch with a NaN member", flush=True)
    script, shape = "diff_3D_linstep_multixpu", (24, 20, 18)
    _, Hnan, D = _nan_ic(orc, script, shape)
    spec = pkg.SCRIPTS[script]
    p = pkg.make_params(spec, *shape, arith=1)
    H0 = pkg.initial_condition(spec, *shape)
    nt = 3
    with pkg.Batch([p, p, p]) as b:
        for m, h in enumerate((H0, Hnan, H0)):
            b.set(m, "H", h)
            b.set(m, "Hold", h)
            b.set(m, "D", D)
        it, er, lat, st, total = b.solve(nt, raise_nan=False)
        assert list(st) == [0, pkg._abi.ERR_NAN, 0]
        assert it[1, 0] == p.nout and np.isnan(er[1, 0]) and np.all(it[1, 1:] == 0)
        good = [b.get(0, "H"), b.get(2, "H")]
        with pytest.raises(pkg.PTDError) as ei:  # member 1 still holds NaN
            b.solve(1)
        assert ei.value.status == pkg._abi.ERR_NAN
    with pkg.Context(p) as ctx:
        ctx.set("H", H0)
        ctx.set("Hold", H0)
        ctx.set("D", D)
        its, errs = ctx.solve(nt)
        Hs = ctx.get("H")
    for m in (0, 2):
        assert np.array_equal(it[m], its) and np.array_equal(er[m], errs)
    assert np.array_equal(good[0], Hs) and np.array_equal(good[1], Hs)


def test_nan_in_the_mixed_solve(pkg, orc):
    print("[gpu] mixed-precision solve from a NaN initial condition", flush=True)
    script, shape = "diff_3D_nonlin_multixpu", (26, 22, 20)
    _, Hnan, _ = _nan_ic(orc, script, shape)
    p = pkg.make_params(pkg.SCRIPTS[script], *shape, precision=pkg.MIXED, arith=1)
    with pkg.Context(p) as ctx:
        ctx.set("H", Hnan)
        ctx.set("Hold", Hnan)
        it, rms, mx = ctx.pt_solve(raise_nan=False)
        with pytest.raises(pkg.PTDError) as ei:
            ctx.pt_solve()
    assert ei.value.status == pkg._abi.ERR_NAN
    assert it == 0 and np.isnan(rms)
