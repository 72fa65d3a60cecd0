"""ptd_pt_solve_noalloc (the PT solve without allocation, on caller-owned device buffers) against
ptd_pt_solve on a context holding the same fields: equal iteration counts and errors, and bit-identical H
and fluxes, over two calls in a row on one device, the second on a smaller grid (the scratch array keeps
the first call's partials). Device buffers come from torch (test plumbing)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _device(pkg):
    lib = pkg.load()
    if lib.ptd_device_check(0, None, 0) != 0:
        pytest.fail(f"no gfx950 device: {lib.ptd_last_error()}")


def _dev(a, dtype):
    """A column-major host array as a flat device tensor (the reference layout's memory order)."""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel(order="F"))).to("cuda")


def _host(t, shape):
    return t.cpu().numpy().astype(np.float64).reshape(shape, order="F")


def _fields(pkg, script, shape, ic):
    spec = pkg.SCRIPTS[script]
    H0 = pkg.initial_condition(spec, *shape, kind=ic)
    D = pkg.step_field(spec, *shape) if spec["dcase"] == pkg.STEP else None
    return spec, H0, D


def _noalloc_solve(pkg, params, H0, D, parity0, stream=None):
    """One ptd_pt_solve_noalloc call from H = Hold = H0 and zero fluxes, with the iterate placed in buffer
    `parity0` (the other H buffer also holds H0: both carry the boundary values)."""
    lib = pkg.load()
    dt = np.float32 if params.precision == pkg.FP32 else np.float64
    nd = params.ndim
    Hb = [_dev(H0, dt), _dev(H0, dt)]
    Hold = _dev(H0, dt)
    Dd = _dev(D, dt) if D is not None else None
    qshapes = [pkg.field_shape(params, f) for f in (3, 4, 5)[:nd]]
    q = [[_dev(np.zeros(s), dt) for _ in range(2)] for s in qshapes]
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    qp = [ptr(q[a][b]) if a < nd else None for a in range(3) for b in range(2)]
    parity = ctypes.c_int32(parity0)
    it, rms, mx = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
    torch.cuda.synchronize()
    st = lib.ptd_pt_solve_noalloc(ctypes.byref(params), 0, stream, ptr(Hb[0]), ptr(Hb[1]), ptr(Hold), ptr(Dd),
                                  *qp, ctypes.byref(parity), ctypes.byref(it), ctypes.byref(rms), ctypes.byref(mx))
    torch.cuda.synchronize()
    assert st == 0, lib.ptd_last_error()
    p = parity.value
    H = _host(Hb[p], H0.shape)
    qs = [_host(q[a][p], qshapes[a]) for a in range(nd)]
    return it.value, rms.value, mx.value, p, H, qs


def _context_solve(pkg, params, H0, D):
    with pkg.Context(params) as ctx:
        ctx.set("H", H0)
        ctx.set("Hold", H0)
        if D is not None:
            ctx.set("D", D)
        it, rms, mx = ctx.pt_solve()
        return it, rms, mx, ctx.get("H"), [ctx.get(f) for f in ("qHx", "qHy", "qHz")[:params.ndim]]


# two calls in a row, the second on a smaller grid
SEQUENCES = [
    [("diff_3D_nonlin_multixpu", (45, 38, 29), "config", 0, 0), ("diff_3D_linstep_multixpu", (20, 18, 16), None, 0, 1)],
    [("diff_3D_lin_multixpu", (70, 33, 40), None, 1, 1), ("diff_2D_nonlin", (64, 48), None, 1, 0)],
    [("diff_2D_linstep", (130, 90), None, 0, 1), ("diff_2D_lin", (40, 33), None, 1, 0)],
]


@pytest.mark.parametrize("seq", range(len(SEQUENCES)))
def test_noalloc_solve_equals_context_solve(seq, pkg):
    for script, shape, ic, arith, parity0 in SEQUENCES[seq]:
        print(f"[gpu] noalloc {script} {shape} arith={arith} parity={parity0}", flush=True)
        spec, H0, D = _fields(pkg, script, shape, ic)
        params = pkg.make_params(spec, *shape, arith=arith)
        it, rms, mx, p, H, qs = _noalloc_solve(pkg, params, H0, D, parity0)
        rit, rrms, rmx, rH, rq = _context_solve(pkg, params, H0, D)
        assert it == rit and it > 0 and it % params.nout == 0
        assert rms == rrms and mx == rmx
        assert p in (0, 1)
        assert np.array_equal(H, rH)
        for a, b in zip(qs, rq):
            assert np.array_equal(a, b)


def test_noalloc_solve_on_a_stream_fp32(pkg):
    """float32 fields on a caller stream: equal to the context's float32 solve."""
    script, shape = "diff_3D_nonlin_multixpu", (33, 30, 27)
    spec, H0, D = _fields(pkg, script, shape, "config")
    params = pkg.make_params(spec, *shape, precision=pkg.FP32, arith=1)
    params.itmax = 200  # float32 residuals stall above 1e-8: a bounded solve with checks
    s = torch.cuda.Stream()
    it, rms, mx, p, H, qs = _noalloc_solve(pkg, params, H0, D, 0, stream=ctypes.c_void_p(s.cuda_stream))
    rit, rrms, rmx, rH, rq = _context_solve(pkg, params, H0, D)
    assert it == rit == 200 and rms == rrms and mx == rmx
    assert np.array_equal(H, rH)
    for a, b in zip(qs, rq):
        assert np.array_equal(a, b)


def test_noalloc_refuses_what_it_cannot_run(pkg):
    """Grids whose checks need more partials than the scratch array holds, 1D, the perf form and the mixed
    precision are refused with PTD_ERR_ARG before any launch."""
    lib = pkg.load()
    one = torch.zeros(1, dtype=torch.float64, device="cuda")
    ptr = ctypes.c_void_p(one.data_ptr())
    parity = ctypes.c_int32(0)
    it, rms, mx = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
    cases = [("diff_3D_nonlin_multixpu", (1400, 1400, 1400), {}), ("diff_1D_lin", (64,), {}),
             ("diff_3D_nonlin_perf", (32, 32, 32), {}),
             ("diff_3D_nonlin_multixpu", (32, 32, 32), {"precision": pkg.MIXED})]
    for script, shape, kw in cases:
        params = pkg.make_params(pkg.SCRIPTS[script], *shape, **kw)
        st = lib.ptd_pt_solve_noalloc(ctypes.byref(params), 0, None, ptr, ptr, ptr, None, ptr, ptr, ptr, ptr,
                                      ptr, ptr, ctypes.byref(parity), ctypes.byref(it), ctypes.byref(rms),
                                      ctypes.byref(mx))
        assert st == 1, (script, shape, st)
    params = pkg.make_params(pkg.SCRIPTS["diff_3D_nonlin_multixpu"], 1400, 1400, 1400)
    lib.ptd_pt_solve_noalloc(ctypes.byref(params), 0, None, ptr, ptr, ptr, None, ptr, ptr, ptr, ptr, ptr, ptr,
                             ctypes.byref(parity), ctypes.byref(it), ctypes.byref(rms), ctypes.byref(mx))
    assert b"too large" in lib.ptd_last_error() and b"131072" in lib.ptd_last_error()
