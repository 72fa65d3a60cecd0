"""GPU parity: libptd (HIP, gfx950) through the C ABI against the CPU oracle and the reference goldens.

Tolerances (float64): the device kernels evaluate every expression in the reference's operation order
with IEEE division/sqrt and no contraction, so results are expected to be bit-identical to the oracle;
the tests assert a relative max-norm difference <= 1e-12 (tighter than the 1e-10 target) and report
the number of bit-identical entries. Iteration counts must be equal. Norms are reduced in a different
order than the oracle: errors agree to 1e-12 relative.
float32: relative max-norm difference to the float64 oracle <= 1e-4 after short runs.
"""
import numpy as np
import pytest

from conftest import FIXTURE_SCRIPTS, isapprox_norm, sample

pytestmark = pytest.mark.gpu

RTOL64 = 1e-12


def _relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _device(pkg):
    lib = pkg.load()
    st = lib.ptd_device_check(0, None, 0)
    if st != 0:
        pytest.fail(f"no gfx950 device: {lib.ptd_last_error()}")


# --------------------------------------------------------------------------------------------------
# the reference's reftests, run through the GPU path
# --------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("fixture", sorted(FIXTURE_SCRIPTS))
def test_gpu_reproduces_reftest_and_oracle(fixture, pkg, orc, reftests):
    """The script run on the GPU reproduces the reftest; fed the oracle's own H0 and D it reproduces the
    oracle bit for bit, iteration counts included."""
    script, n = FIXTURE_SCRIPTS[fixture]
    print(f"[gpu] reftest {fixture} n={n}", flush=True)
    res = pkg.run_script(script, n, n, n)
    nd = len(res.H.shape)
    Hs, inds = sample(res.H, n, nd)
    assert np.array_equal(res.coords[0][inds], reftests[fixture]["X"])
    assert isapprox_norm(Hs, reftests[fixture]["H"])
    assert np.linalg.norm(Hs - reftests[fixture]["H"]) / np.linalg.norm(reftests[fixture]["H"]) <= 1e-13
    ref = orc.run_script(script, n, n, n)
    same = pkg.run_script(script, n, n, n, H0=ref["H0"], D=ref["D"])
    assert list(same.iters) == list(ref["iters"])
    assert np.array_equal(same.H, ref["H"]), f"rel {_relmax(same.H, ref['H']):.2e}"
    np.testing.assert_allclose(same.errs, ref["errs"], rtol=1e-12)


# --------------------------------------------------------------------------------------------------
# kernel-level entry points (one reference kernel each) on random states
# --------------------------------------------------------------------------------------------------

def _random_state(pkg, orc, script, shape, seed):
    rng = np.random.default_rng(seed)
    spec = pkg.SCRIPTS[script]
    nx, ny, nz = (list(shape) + [1, 1])[:3]
    params = pkg.make_params(spec, nx, ny, nz)
    case = orc.make_case(orc.SCRIPTS[script], nx, ny, nz)
    # positive, mildly varying states: the nonlinear PT sweep stays in its domain (H > 0)
    H = np.asfortranarray(rng.uniform(1.0, 1.2, size=shape))
    Hold = np.asfortranarray(rng.uniform(1.0, 1.2, size=shape))
    D = np.asfortranarray(np.where(rng.uniform(size=shape) < 0.5, 1e-4, 1.0)) if spec["dcase"] == pkg.STEP else None
    q = {}
    for name in ("qHx", "qHy", "qHz")[:len(shape)]:
        q[name] = np.asfortranarray(0.05 * rng.normal(size=pkg.field_shape(params, {"qHx": 3, "qHy": 4, "qHz": 5}[name])))
    return params, case, H, Hold, D, q


KERNEL_CASES = [
    ("diff_1D_lin", (23,)), ("diff_1D_linstep", (23,)), ("diff_1D_nonlin", (23,)),
    ("diff_2D_lin", (19, 13)), ("diff_2D_linstep", (19, 13)), ("diff_2D_nonlin", (19, 13)),
    ("diff_2D_nonlin_perf", (19, 13)),
    ("diff_3D_lin_multixpu", (11, 9, 7)), ("diff_3D_linstep_multixpu", (11, 9, 7)),
    ("diff_3D_nonlin_multixpu", (11, 9, 7)), ("diff_3D_nonlin_perf", (11, 9, 7)),
]


@pytest.mark.parametrize("script,shape", KERNEL_CASES)
def test_kernel_level_entry_points(script, shape, pkg, orc):
    print(f"[gpu] kernels {script} {shape}", flush=True)
    params, case, H, Hold, D, q = _random_state(pkg, orc, script, shape, seed=len(script))
    nd = len(shape)
    qn = ("qHx", "qHy", "qHz")[:nd]
    with pkg.Context(params) as ctx:
        ctx.set("H", H)
        ctx.set("Hold", Hold)
        for k, v in q.items():
            ctx.set(k, v)
        dtau = None
        if D is not None:
            ctx.set("D", D)
            ctx.compute_iter_params()
            dtau = np.zeros(pkg.field_shape(params, 10), order="F")
            orc.lib().ptdo_iter_params(orc.ctypes.byref(case), orc._p(D), orc._p(dtau))
            assert np.array_equal(ctx.get("dtau_rho"), dtau)
        # compute_flux!
        ctx.compute_flux()
        qo = {k: v.copy(order="F") for k, v in q.items()}
        q2o = {k: np.zeros_like(v) for k, v in q.items()}
        args = [qo.get(k) for k in ("qHx", "qHy", "qHz")] + [q2o.get(k) for k in ("qHx", "qHy", "qHz")]
        orc.lib().ptdo_flux(orc.ctypes.byref(case), orc._p(H), orc._p(D), *[orc._p(a) for a in args])
        for k in qn:
            assert np.array_equal(ctx.get(k), qo[k]), k
            if params.form == pkg.PLAIN:
                assert np.array_equal(ctx.get(k + "2"), q2o[k]), k + "2"
        # compute_update!
        ctx.compute_update()
        Hr = H.copy(order="F")
        orc.lib().ptdo_update(orc.ctypes.byref(case), orc._p(Hr), orc._p(Hold), orc._p(dtau),
                             *[orc._p(qo.get(k)) for k in ("qHx", "qHy", "qHz")])
        assert np.array_equal(ctx.get("H"), Hr)
        # compute_flux_res! + check_res!
        ctx.compute_flux_res()
        orc.lib().ptdo_flux_res(orc.ctypes.byref(case), orc._p(Hr), orc._p(D),
                               *[orc._p(q2o.get(k)) for k in ("qHx", "qHy", "qHz")])
        for k in qn:
            assert np.array_equal(ctx.get(k + "2"), q2o[k])
        rms, mx = ctx.check_res()
        resh = np.zeros(pkg.field_shape(params, 9), order="F")
        emax = orc.ctypes.c_double()
        ref_rms = orc.lib().ptdo_check_res(orc.ctypes.byref(case), orc._p(Hr), orc._p(Hold),
                                          *[orc._p(q2o.get(k)) for k in ("qHx", "qHy", "qHz")],
                                          orc._p(resh), orc.ctypes.byref(emax))
        assert np.array_equal(ctx.get("ResH"), resh)
        assert rms == pytest.approx(ref_rms, rel=1e-12) and mx == emax.value
        # assign!
        ctx.assign()
        assert np.array_equal(ctx.get("Hold"), Hr)


# --------------------------------------------------------------------------------------------------
# fused 3D sweep on ragged shapes: several x / y tiles, partial tiles, several z chunks
# --------------------------------------------------------------------------------------------------

FUSED_SHAPES = [(37, 29, 21), (70, 33, 18), (131, 19, 47), (3, 3, 3), (64, 16, 5)]
FUSED_SCRIPTS = ["diff_3D_lin_multixpu", "diff_3D_linstep_multixpu", "diff_3D_nonlin_multixpu",
                 "diff_3D_nonlin_perf"]


@pytest.mark.parametrize("shape", FUSED_SHAPES)
@pytest.mark.parametrize("script", FUSED_SCRIPTS)
def test_fused_sweeps_match_oracle(script, shape, pkg, orc, monkeypatch):
    print(f"[gpu] fused {script} {shape}", flush=True)
    monkeypatch.setenv("PTD_TZ", "4")  # several z chunks per column
    params, case, H, Hold, D, q = _random_state(pkg, orc, script, shape, seed=sum(shape))
    niter = 5
    with pkg.Context(params) as ctx:
        ctx.set("H", H)
        ctx.set("Hold", Hold)
        for k, v in q.items():
            ctx.set(k, v)
        dtau = None
        if D is not None:
            ctx.set("D", D)
            dtau = np.zeros(pkg.field_shape(params, 10), order="F")
            orc.lib().ptdo_iter_params(orc.ctypes.byref(case), orc._p(D), orc._p(dtau))
        ctx.iterate(niter)
        Hg = ctx.get("H")
        qg = {k: ctx.get(k) for k in q}
    Hr = H.copy(order="F")
    qr = {k: v.copy(order="F") for k, v in q.items()}
    orc.lib().ptdo_iterate(orc.ctypes.byref(case), orc._p(Hr), orc._p(Hold), orc._p(D), orc._p(dtau),
                          orc._p(qr["qHx"]), orc._p(qr["qHy"]), orc._p(qr["qHz"]), niter)
    assert np.all(np.isfinite(Hr))
    assert np.array_equal(Hg, Hr), f"H: rel {_relmax(Hg, Hr):.2e}"
    for k in q:
        assert np.array_equal(qg[k], qr[k]), f"{k}: rel {_relmax(qg[k], qr[k]):.2e}"


@pytest.mark.parametrize("script", ["diff_3D_nonlin_multixpu", "diff_3D_linstep_multixpu", "diff_3D_lin_multixpu"])
def test_fused_solve_with_checks_matches_oracle(script, pkg, orc):
    """Converging solve at 45x38x29: residual checks fused into the sweep (plain form)."""
    shape = (45, 38, 29)
    print(f"[gpu] solve {script} {shape}", flush=True)
    ref = orc.run_script(script, *shape)
    res = pkg.run_script(script, *shape, H0=ref["H0"], D=ref["D"])
    assert list(res.iters) == list(ref["iters"])
    assert np.array_equal(res.H, ref["H"]), f"rel {_relmax(res.H, ref['H']):.2e}"
    np.testing.assert_allclose(res.errs, ref["errs"], rtol=1e-12)


def test_perf_form_solve_with_checks(pkg, orc):
    """Perf-form expressions with a converging loop (nout=10): residual from the post-update H."""
    ref = orc.run_script("diff_3D_nonlin_perf", 33, 30, 27, nout=10, itmax=100000)
    res = pkg.run_script("diff_3D_nonlin_perf", 33, 30, 27, nout=10, itMax=1e5, H0=ref["H0"])
    assert list(res.iters) == list(ref["iters"])
    assert np.array_equal(res.H, ref["H"])


def test_max_norm_stopping(pkg, orc):
    res = pkg.run_script("diff_2D_nonlin", 48, 48, stop_norm=1)
    assert all(e <= 1e-8 for e in res.errs)
    base = pkg.run_script("diff_2D_nonlin", 48, 48)
    assert res.ittot >= base.ittot  # max|r| >= rms(r)


# --------------------------------------------------------------------------------------------------
# float32 sweeps
# --------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("script", ["diff_3D_nonlin_perf", "diff_2D_nonlin_perf", "diff_3D_linstep_multixpu"])
def test_fp32_sweeps_close_to_fp64_oracle(script, pkg, orc):
    n = 32
    over = {"itMax": 100, "nout": 2000} if "linstep" in script else {"itMax": 100}
    res = pkg.run_script(script, n, n, n, precision=pkg.FP32, ic="config" if "nonlin" in script else None, **over)
    spec = dict(orc.SCRIPTS[script])
    spec.update(itmax=100, nout=2000)
    case = orc.make_case(spec, n, n, n)
    H0 = pkg.initial_condition(pkg.SCRIPTS[script], n, n, n, kind="config" if "nonlin" in script else None)
    H, Hold = H0.copy(order="F"), H0.copy(order="F")
    D = pkg.step_field(pkg.SCRIPTS[script], n, n, n) if "linstep" in script else None
    orc.solve(case, H, Hold, D)
    assert _relmax(res.H, H) <= 1e-4


# --------------------------------------------------------------------------------------------------
# full-size properties (511^3, 1 x fused sweeps)
# --------------------------------------------------------------------------------------------------

def test_full_size_sweeps_properties(pkg, orc):
    """511^3 nonlinear, config IC: boundaries unchanged, bitwise x<->y symmetry, finite; and a 2-plane
    z-slab comparison against the oracle on a 255^3 run."""
    n = 511
    print(f"[gpu] full size {n}^3", flush=True)
    spec = pkg.SCRIPTS["diff_3D_nonlin_perf"]
    params = pkg.make_params(spec, n, n, n)
    H0 = pkg.initial_condition(spec, n, n, n, kind="config")
    with pkg.Context(params) as ctx:
        ctx.set("H", H0)
        ctx.set("Hold", H0)
        ctx.iterate(20)
        H = ctx.get("H")
    assert np.all(np.isfinite(H))
    for d in range(3):
        for edge in (0, n - 1):
            sl = [slice(None)] * 3
            sl[d] = edge
            assert np.array_equal(H[tuple(sl)], H0[tuple(sl)])
    assert np.array_equal(H, np.swapaxes(H, 0, 1))  # (dx-term + dy-term) + dz-term: x<->y is exact
    del H, H0
    n = 255
    params = pkg.make_params(spec, n, n, n)
    H0 = pkg.initial_condition(spec, n, n, n, kind="config")
    with pkg.Context(params) as ctx:
        ctx.set("H", H0)
        ctx.set("Hold", H0)
        ctx.iterate(3)
        Hg = ctx.get("H")
    case = orc.make_case(orc.SCRIPTS["diff_3D_nonlin_perf"], n, n, n)
    Hr, Ho = H0.copy(order="F"), H0.copy(order="F")
    z = [np.zeros(s, order="F") for s in ((n - 1, n - 2, n - 2), (n - 2, n - 1, n - 2), (n - 2, n - 2, n - 1))]
    orc.lib().ptdo_iterate(orc.ctypes.byref(case), orc._p(Hr), orc._p(Ho), None, None,
                          orc._p(z[0]), orc._p(z[1]), orc._p(z[2]), 3)
    assert np.array_equal(Hg, Hr)


# --------------------------------------------------------------------------------------------------
# FAST arithmetic (rearranged update rules): close to the oracle, same converged solutions
# --------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("shape", [(37, 29, 21), (131, 19, 47)])
@pytest.mark.parametrize("script", FUSED_SCRIPTS)
def test_fast_sweeps_close_to_oracle(script, shape, pkg, orc, monkeypatch):
    """5 FAST sweeps from a random state: relative max-norm difference <= 1e-13 (a few ulp per sweep)."""
    print(f"[gpu] fast {script} {shape}", flush=True)
    monkeypatch.setenv("PTD_TZ", "4")
    params, case, H, Hold, D, q = _random_state(pkg, orc, script, shape, seed=sum(shape) + 1)
    params.arith = 1
    with pkg.Context(params) as ctx:
        ctx.set("H", H)
        ctx.set("Hold", Hold)
        for k, v in q.items():
            ctx.set(k, v)
        dtau = None
        if D is not None:
            ctx.set("D", D)
            dtau = np.zeros(pkg.field_shape(params, 10), order="F")
            orc.lib().ptdo_iter_params(orc.ctypes.byref(case), orc._p(D), orc._p(dtau))
        ctx.iterate(5)
        Hg = ctx.get("H")
        qg = {k: ctx.get(k) for k in q}
    Hr = H.copy(order="F")
    qr = {k: v.copy(order="F") for k, v in q.items()}
    orc.lib().ptdo_iterate(orc.ctypes.byref(case), orc._p(Hr), orc._p(Hold), orc._p(D), orc._p(dtau),
                           orc._p(qr["qHx"]), orc._p(qr["qHy"]), orc._p(qr["qHz"]), 5)
    assert np.all(np.isfinite(Hr))
    assert _relmax(Hg, Hr) <= 1e-13
    for k in q:
        assert _relmax(qg[k], qr[k]) <= 1e-12, k


@pytest.mark.parametrize("script", ["diff_3D_nonlin_multixpu", "diff_3D_linstep_multixpu", "diff_3D_lin_multixpu",
                                    "diff_3D_nonlin_perf"])
def test_fast_solve_matches_oracle(script, pkg, orc):
    """Converging solve with FAST sweeps: same per-step iteration counts (within one nout) and final H
    within 1e-10 relative max-norm of the oracle."""
    shape = (45, 38, 29)
    print(f"[gpu] fast solve {script} {shape}", flush=True)
    over = dict(nout=10, itmax=100000) if "perf" in script else {}
    ref = orc.run_script(script, *shape, **over)
    over = dict(nout=10, itMax=1e5) if "perf" in script else {}
    res = pkg.run_script(script, *shape, H0=ref["H0"], D=ref["D"], arith=1, **over)
    assert all(abs(int(a) - int(b)) <= 10 for a, b in zip(res.iters, ref["iters"])), (list(res.iters), list(ref["iters"]))
    assert _relmax(res.H, ref["H"]) <= 1e-10


# --------------------------------------------------------------------------------------------------
# two-level (temporally blocked) sweep: pairs of iterations in one pass
# --------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("niter", [2, 6, 7])
@pytest.mark.parametrize("shape", [(37, 29, 21), (131, 19, 47), (64, 16, 5), (3, 3, 3)])
@pytest.mark.parametrize("script", FUSED_SCRIPTS)
def test_two_level_sweeps_match_oracle(script, shape, niter, pkg, orc, monkeypatch):
    """EXACT arithmetic: pairs through the two-level sweep (plus one single sweep for odd counts) give
    the oracle's values bit for bit, with 4-plane z chunks."""
    print(f"[gpu] t2 {script} {shape} niter={niter}", flush=True)
    monkeypatch.setenv("PTD_TZ", "4")
    monkeypatch.setenv("PTD_T2", "1")
    params, case, H, Hold, D, q = _random_state(pkg, orc, script, shape, seed=sum(shape) + niter)
    with pkg.Context(params) as ctx:
        ctx.set("H", H)
        ctx.set("Hold", Hold)
        for k, v in q.items():
            ctx.set(k, v)
        dtau = None
        if D is not None:
            ctx.set("D", D)
            dtau = np.zeros(pkg.field_shape(params, 10), order="F")
            orc.lib().ptdo_iter_params(orc.ctypes.byref(case), orc._p(D), orc._p(dtau))
        ctx.iterate(niter)
        Hg = ctx.get("H")
        qg = {k: ctx.get(k) for k in q}
    Hr = H.copy(order="F")
    qr = {k: v.copy(order="F") for k, v in q.items()}
    orc.lib().ptdo_iterate(orc.ctypes.byref(case), orc._p(Hr), orc._p(Hold), orc._p(D), orc._p(dtau),
                           orc._p(qr["qHx"]), orc._p(qr["qHy"]), orc._p(qr["qHz"]), niter)
    assert np.all(np.isfinite(Hr))
    assert np.array_equal(Hg, Hr), f"H: rel {_relmax(Hg, Hr):.2e}"
    for k in q:
        assert np.array_equal(qg[k], qr[k]), f"{k}: rel {_relmax(qg[k], qr[k]):.2e}"


@pytest.mark.parametrize("script", FUSED_SCRIPTS)
def test_two_level_equals_single_sweeps_fast(script, pkg, orc, monkeypatch):
    """FAST arithmetic: 8 iterations as 4 two-level passes equal 8 single sweeps bit for bit."""
    shape = (70, 33, 18)
    out = []
    for t2 in ("1", "0"):
        monkeypatch.setenv("PTD_T2", t2)
        params, case, H, Hold, D, q = _random_state(pkg, orc, script, shape, seed=7)
        params.arith = 1
        with pkg.Context(params) as ctx:
            ctx.set("H", H)
            ctx.set("Hold", Hold)
            for k, v in q.items():
                ctx.set(k, v)
            if D is not None:
                ctx.set("D", D)
            ctx.iterate(8)
            out.append((ctx.get("H"), [ctx.get(k) for k in q]))
    assert np.array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)


# Two-level 2D sweeps run on grids with at least 1024 tiles of 60 x 16 cells; (1000, 990) has 17 x 62.
T2D_SCRIPTS = ["diff_2D_lin", "diff_2D_linstep", "diff_2D_nonlin", "diff_2D_nonlin_perf"]
T2D_SHAPE = (1000, 990)


@pytest.mark.parametrize("script", T2D_SCRIPTS)
def test_two_level_2d_sweeps_match_oracle(script, pkg, orc, monkeypatch):
    """EXACT arithmetic, 2D: three two-level passes plus one single sweep give the oracle's values bit for
    bit (H and both fluxes)."""
    niter = 7
    print(f"[gpu] t2 2D {script} {T2D_SHAPE} niter={niter}", flush=True)
    monkeypatch.setenv("PTD_T2", "1")
    monkeypatch.setenv("PTD_LDS2D", "0")  # the two-level sweeps, not the LDS-resident iterations
    params, case, H, Hold, D, q = _random_state(pkg, orc, script, T2D_SHAPE, seed=11)
    with pkg.Context(params) as ctx:
        ctx.set("H", H)
        ctx.set("Hold", Hold)
        for k, v in q.items():
            ctx.set(k, v)
        dtau = None
        if D is not None:
            ctx.set("D", D)
            dtau = np.zeros(pkg.field_shape(params, 10), order="F")
            orc.lib().ptdo_iter_params(orc.ctypes.byref(case), orc._p(D), orc._p(dtau))
        ctx.iterate(niter)
        Hg = ctx.get("H")
        qg = {k: ctx.get(k) for k in q}
    Hr = H.copy(order="F")
    qr = {k: v.copy(order="F") for k, v in q.items()}
    orc.lib().ptdo_iterate(orc.ctypes.byref(case), orc._p(Hr), orc._p(Hold), orc._p(D), orc._p(dtau),
                           orc._p(qr["qHx"]), orc._p(qr["qHy"]), None, niter)
    assert np.all(np.isfinite(Hr))
    assert np.array_equal(Hg, Hr), f"H: rel {_relmax(Hg, Hr):.2e}"
    for k in q:
        assert np.array_equal(qg[k], qr[k]), f"{k}: rel {_relmax(qg[k], qr[k]):.2e}"


@pytest.mark.parametrize("script", ["diff_2D_lin", "diff_2D_linstep", "diff_2D_nonlin"])
@pytest.mark.parametrize("precision", [0, 1], ids=["f64", "f32"])
def test_two_level_2d_converged_step_equals_single_level(script, precision, pkg, orc, monkeypatch):
    """FAST arithmetic, 2D: one converged physical step with two-level sweeps (and their checking variant)
    takes the same iterations as with single sweeps and ends on the same bits; the residuals agree to
    summation order."""
    print(f"[gpu] t2 2D converged step {script} precision={precision}", flush=True)
    spec = pkg.SCRIPTS[script]
    out = []
    monkeypatch.setenv("PTD_LDS2D", "0")
    for t2 in ("1", "0"):
        monkeypatch.setenv("PTD_T2", t2)
        params = pkg.make_params(spec, *T2D_SHAPE, precision=precision, arith=1)
        if precision == 1:
            params.itmax = 200  # float32 residuals stall above 1e-8: a bounded run with checks
        H0 = pkg.initial_condition(spec, *T2D_SHAPE)
        with pkg.Context(params) as ctx:
            ctx.set("H", H0)
            ctx.set("Hold", H0)
            if spec["dcase"] == pkg.STEP:
                ctx.set("D", pkg.step_field(spec, *T2D_SHAPE))
            it, rms, mx = ctx.pt_solve()
            out.append((it, rms, ctx.get("H"), [ctx.get(k) for k in ("qHx", "qHy")]))
    assert out[0][0] == out[1][0] and out[0][0] > 0
    rtol = 1e-12 if precision == 0 else 1e-5  # per-thread partial sums in the sweep's precision
    assert abs(out[0][1] - out[1][1]) <= rtol * abs(out[1][1]) + 1e-300
    assert np.array_equal(out[0][2], out[1][2])
    for a, b in zip(out[0][3], out[1][3]):
        assert np.array_equal(a, b)


# --------------------------------------------------------------------------------------------------
# mixed precision: float32 correction sweeps inside float64 defect correction
# --------------------------------------------------------------------------------------------------

MIXED_CASES = [("diff_3D_nonlin_multixpu", "config"), ("diff_3D_linstep_multixpu", None),
               ("diff_3D_lin_multixpu", None), ("diff_3D_nonlin_multixpu", None)]


@pytest.mark.parametrize("arith", [0, 1], ids=["exact", "fast"])
@pytest.mark.parametrize("script,ic", MIXED_CASES)
def test_mixed_solve_matches_fp64_oracle(script, ic, arith, pkg, orc):
    """5 physical steps at 45x38x29: every step ends with the float64 residual <= tol, and the final H
    agrees with the float64 oracle's (both converged to tol 1e-8) to 1e-7 relative max-norm. arith selects
    the float64 passes of the defect correction (the benchmark runs FAST)."""
    shape = (45, 38, 29)
    print(f"[gpu] mixed {script} ic={ic} arith={arith}", flush=True)
    spec = dict(orc.SCRIPTS[script])
    case = orc.make_case(spec, *shape)
    H0 = pkg.initial_condition(pkg.SCRIPTS[script], *shape, kind=ic)
    D = pkg.step_field(pkg.SCRIPTS[script], *shape) if "linstep" in script else None
    Hr, Hold_r = H0.copy(order="F"), H0.copy(order="F")
    it_ref, _, st = orc.solve(case, Hr, Hold_r, D)
    assert st == 0
    params = pkg.make_params(pkg.SCRIPTS[script], *shape, precision=pkg.MIXED, arith=arith)
    with pkg.Context(params) as ctx:
        ctx.set("H", H0)
        ctx.set("Hold", H0)
        if D is not None:
            ctx.set("D", D)
        for step in range(case.nt):
            n, rms, mx = ctx.pt_solve()
            assert rms <= params.tol, (step, rms)
            # the consistent float64 residual of the returned iterate, recomputed by the kernel-level API
            ctx.compute_flux_res()
            rms2, _ = ctx.check_res()
            assert rms2 <= params.tol
            assert ctx.outer_iterations >= 1
            print(f"[gpu]   step {step}: {n} float32 sweeps, {ctx.outer_iterations} corrections, "
                  f"rms {rms:.2e} (oracle {it_ref[step]} iterations)", flush=True)
            ctx.assign()
        Hm = ctx.get("H")
    assert _relmax(Hm, Hr) <= 1e-7, _relmax(Hm, Hr)


def test_fp32_reference_ic_no_underflow_nan(pkg, orc):
    """float32 sweeps on the reference IC (H ~ 1e-31 near the corners, H^3 below the float32 range):
    finite, and within 1e-4 of the float64 oracle after 2 x 100 iterations."""
    n = 32
    res = pkg.run_script("diff_3D_nonlin_perf", n, n, n, precision=pkg.FP32, itMax=100, arith=1)
    ref = orc.run_script("diff_3D_nonlin_perf", n, n, n, itmax=100)
    assert np.all(np.isfinite(res.H))
    assert _relmax(res.H, ref["H"]) <= 1e-4


@pytest.mark.parametrize("script,shape", [("diff_3D_nonlin_multixpu", (29, 27, 25)), ("diff_2D_linstep", (50, 44)),
                                          ("diff_1D_nonlin", (64,))])
def test_itmax_between_checks_matches_oracle(script, shape, pkg, orc):
    """itMax that is not a multiple of nout: each step runs itMax iterations, the last check at the
    largest multiple of nout below it sets err, and the loop ends on iter >= itMax (the `while err > tol &&
    iter < itMax` of scripts/diff_3D/diff_3D_nonlin_multixpu.jl:104), unless a check already met tol (the
    3D case converges at 20 after its first step); 25 iterations are two-level pairs plus an odd single
    sweep in the fused paths."""
    print(f"[gpu] itmax 25 {script} {shape}", flush=True)
    ref = orc.run_script(script, *shape, itmax=25)
    res = pkg.run_script(script, *shape, itMax=25, H0=ref["H0"], D=ref["D"])
    assert list(res.iters) == list(ref["iters"]) and 25 in list(ref["iters"])
    assert np.array_equal(res.H, ref["H"]), f"rel {_relmax(res.H, ref['H']):.2e}"
    np.testing.assert_allclose(res.errs, ref["errs"], rtol=1e-12)


def test_mixed_solve_stops_at_itmax(pkg, orc):
    """itMax bounds the float32 sweeps of a mixed-precision PT solve as it bounds the iterations of the
    float64 one: the solve returns PTD_OK with its float64 residual above tol, and the next call continues
    from there."""
    shape = (45, 38, 29)
    print(f"[gpu] mixed solve with itMax 20 {shape}", flush=True)
    spec = pkg.SCRIPTS["diff_3D_nonlin_multixpu"]
    params = pkg.make_params(spec, *shape, precision=pkg.MIXED)
    params.itmax = 20
    H0 = pkg.initial_condition(spec, *shape, kind="config")
    with pkg.Context(params) as ctx:
        ctx.set("H", H0)
        ctx.set("Hold", H0)
        n1, rms1, _ = ctx.pt_solve()
        n2, rms2, _ = ctx.pt_solve()
    assert n1 == 20 and 0 < n2 <= 20
    assert rms1 > params.tol and rms2 < rms1


@pytest.mark.parametrize("arith", [0, 1], ids=["exact", "fast"])
@pytest.mark.parametrize("shape", [(3, 3, 3), (5, 5, 5), (3, 4, 6), (8, 7, 9), (70, 9, 5), (9, 70, 34)])
def test_mixed_solve_small_and_ragged_grids(shape, arith, pkg, orc):
    """The mixed-precision solve on grids down to one inner cell and on thin / ragged boxes (one partial
    x-y tile of the float64 pass, several z chunks): every step's float64 residual <= tol, and the final H
    within 1e-7 of the float64 oracle."""
    print(f"[gpu] mixed small grid {shape} arith={arith}", flush=True)
    script = "diff_3D_nonlin_multixpu"
    case = orc.make_case(orc.SCRIPTS[script], *shape)
    H0 = pkg.initial_condition(pkg.SCRIPTS[script], *shape, kind="config")
    Hr, Hold_r = H0.copy(order="F"), H0.copy(order="F")
    _, _, st = orc.solve(case, Hr, Hold_r, None)
    assert st == 0
    params = pkg.make_params(pkg.SCRIPTS[script], *shape, precision=pkg.MIXED, arith=arith)
    with pkg.Context(params) as ctx:
        ctx.set("H", H0)
        ctx.set("Hold", H0)
        for step in range(case.nt):
            n, rms, mx = ctx.pt_solve()
            assert rms <= params.tol, (step, rms)
            ctx.assign()
        Hm = ctx.get("H")
    assert _relmax(Hm, Hr) <= 1e-7, _relmax(Hm, Hr)


@pytest.mark.parametrize("shape", [(20, 18, 100), (481, 258, 100)])
@pytest.mark.parametrize("script", FUSED_SCRIPTS)
def test_two_level_default_chunks_match_oracle(script, shape, pkg, orc, monkeypatch):
    """EXACT arithmetic with the library's own z chunk depth on tall grids: 98 inner planes are 4-plane
    chunks and a 2-plane one on the narrow grid (too few tiles to fill the GPU otherwise), and 32- or 24-plane
    chunks and a 2-plane one on the wide grid; 6 iterations give the oracle's values bit for bit."""
    print(f"[gpu] default chunks {script} {shape}", flush=True)
    monkeypatch.delenv("PTD_TZ", raising=False)
    params, case, H, Hold, D, q = _random_state(pkg, orc, script, shape, seed=7)
    with pkg.Context(params) as ctx:
        ctx.set("H", H)
        ctx.set("Hold", Hold)
        for k, v in q.items():
            ctx.set(k, v)
        dtau = None
        if D is not None:
            ctx.set("D", D)
            dtau = np.zeros(pkg.field_shape(params, 10), order="F")
            orc.lib().ptdo_iter_params(orc.ctypes.byref(case), orc._p(D), orc._p(dtau))
        ctx.iterate(6)
        Hg = ctx.get("H")
        qg = {k: ctx.get(k) for k in q}
    Hr = H.copy(order="F")
    qr = {k: v.copy(order="F") for k, v in q.items()}
    orc.lib().ptdo_iterate(orc.ctypes.byref(case), orc._p(Hr), orc._p(Hold), orc._p(D), orc._p(dtau),
                           orc._p(qr["qHx"]), orc._p(qr["qHy"]), orc._p(qr["qHz"]), 6)
    assert np.array_equal(Hg, Hr), f"H: rel {_relmax(Hg, Hr):.2e}"
    for k in q:
        assert np.array_equal(qg[k], qr[k]), f"{k}: rel {_relmax(qg[k], qr[k]):.2e}"
