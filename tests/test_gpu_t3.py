"""GPU parity of the three-level 3D sweep (k_fused3d_t3: three PT iterations per HBM pass), against the CPU
oracle and against the one- and two-level sweeps.

 - EXACT arithmetic: H and the three fluxes bit-identical to the oracle's niter iterations (3 = one pass,
   7 = two passes and one single sweep, 8 = two passes and one two-level pass) for the four 3D scripts
   (scripts/diff_3D/*.jl compute_flux! + compute_update!), on ragged grids, the smallest grid and grids
   whose z chunks (4 planes, and the default depth) split the columns;
 - FAST arithmetic: the same bits as the same iterations through the two-level / single sweeps;
 - a converging solve with checks (every nout = 10: three-level passes, then a checking two-level pass):
   the oracle's counts and H.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRIPTS_3D = ["diff_3D_lin_multixpu", "diff_3D_linstep_multixpu", "diff_3D_nonlin_multixpu", "diff_3D_nonlin_perf"]
SHAPES = [(37, 29, 21), (70, 33, 18), (131, 19, 47), (3, 3, 3), (64, 16, 5), (61, 62, 40)]


def _relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _device(pkg):
    lib = pkg.load()
    if lib.ptd_device_check(0, None, 0) != 0:
        pytest.fail(f"no gfx950 device: {lib.ptd_last_error()}")


def _random_state(pkg, orc, script, shape, seed, arith=0, precision=0):
    rng = np.random.default_rng(seed)
    spec = pkg.SCRIPTS[script]
    params = pkg.make_params(spec, *shape, arith=arith, precision=precision)
    case = orc.make_case(orc.SCRIPTS[script], *shape)
    H = np.asfortranarray(rng.uniform(1.0, 1.2, size=shape))
    Hold = np.asfortranarray(rng.uniform(1.0, 1.2, size=shape))
    D = np.asfortranarray(np.where(rng.uniform(size=shape) < 0.5, 1e-4, 1.0)) if spec["dcase"] == pkg.STEP else None
    q = {name: np.asfortranarray(0.05 * rng.normal(size=pkg.field_shape(params, fid)))
         for name, fid in (("qHx", 3), ("qHy", 4), ("qHz", 5))}
    return params, case, H, Hold, D, q


def _gpu_iterate(pkg, params, H, Hold, D, q, niter):
    with pkg.Context(params) as ctx:
        ctx.set("H", H)
        ctx.set("Hold", Hold)
        for k, v in q.items():
            ctx.set(k, v)
        if D is not None:
            ctx.set("D", D)
        ctx.iterate(niter)
        return ctx.get("H"), {k: ctx.get(k) for k in q}


def _oracle_iterate(orc, pkg, params, case, H, Hold, D, q, niter):
    dtau = None
    if D is not None:
        dtau = np.zeros(pkg.field_shape(params, 10), order="F")
        orc.lib().ptdo_iter_params(orc.ctypes.byref(case), orc._p(D), orc._p(dtau))
    Hr = H.copy(order="F")
    qr = {k: v.copy(order="F") for k, v in q.items()}
    orc.lib().ptdo_iterate(orc.ctypes.byref(case), orc._p(Hr), orc._p(Hold), orc._p(D), orc._p(dtau),
                           orc._p(qr["qHx"]), orc._p(qr["qHy"]), orc._p(qr["qHz"]), niter)
    return Hr, qr


@pytest.mark.parametrize("tz", ["4", "default"])
@pytest.mark.parametrize("niter", [3, 7, 8])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("script", SCRIPTS_3D)
def test_t3_sweeps_match_oracle(script, shape, niter, tz, pkg, orc, monkeypatch):
    print(f"[gpu] t3 {script} {shape} niter={niter} tz={tz}", flush=True)
    monkeypatch.setenv("PTD_T3", "1")
    if tz == "default":
        monkeypatch.delenv("PTD_TZ", raising=False)
    else:
        monkeypatch.setenv("PTD_TZ", tz)
    params, case, H, Hold, D, q = _random_state(pkg, orc, script, shape, seed=sum(shape) + niter)
    Hg, qg = _gpu_iterate(pkg, params, H, Hold, D, q, niter)
    Hr, qr = _oracle_iterate(orc, pkg, params, case, H, Hold, D, q, niter)
    assert np.all(np.isfinite(Hr))
    assert np.array_equal(Hg, Hr), f"H: rel {_relmax(Hg, Hr):.2e}"
    for k in q:
        assert np.array_equal(qg[k], qr[k]), f"{k}: rel {_relmax(qg[k], qr[k]):.2e}"


@pytest.mark.parametrize("precision", [0, 1], ids=["f64", "f32"])
@pytest.mark.parametrize("script", SCRIPTS_3D)
def test_t3_fast_equals_two_level(script, precision, pkg, orc, monkeypatch):
    shape, niter = (70, 33, 18), 9
    print(f"[gpu] t3 fast vs two-level {script} precision={precision}", flush=True)
    monkeypatch.setenv("PTD_TZ", "4")
    params, case, H, Hold, D, q = _random_state(pkg, orc, script, shape, seed=3, arith=1, precision=precision)
    out = []
    for t3 in ("1", "0"):
        monkeypatch.setenv("PTD_T3", t3)
        out.append(_gpu_iterate(pkg, params, H, Hold, D, q, niter))
    assert np.all(np.isfinite(out[1][0]))
    assert np.array_equal(out[0][0], out[1][0]), f"H: rel {_relmax(out[0][0], out[1][0]):.2e}"
    for k in q:
        assert np.array_equal(out[0][1][k], out[1][1][k]), k


@pytest.mark.parametrize("script", ["diff_3D_nonlin_multixpu", "diff_3D_linstep_multixpu", "diff_3D_lin_multixpu"])
def test_t3_solve_with_checks_matches_oracle(script, pkg, orc, monkeypatch):
    shape = (45, 38, 29)
    print(f"[gpu] t3 solve {script} {shape}", flush=True)
    monkeypatch.setenv("PTD_T3", "1")
    ref = orc.run_script(script, *shape)
    res = pkg.run_script(script, *shape, H0=ref["H0"], D=ref["D"])
    assert list(res.iters) == list(ref["iters"])
    assert np.array_equal(res.H, ref["H"]), f"rel {_relmax(res.H, ref['H']):.2e}"
    np.testing.assert_allclose(res.errs, ref["errs"], rtol=1e-12)


@pytest.mark.parametrize("script", ["diff_3D_lin_multixpu", "diff_3D_linstep_multixpu", "diff_3D_nonlin_perf"])
def test_t3_full_size_equals_two_level(script, pkg, monkeypatch):
    """At 511^3 the three-level passes take the depth of their own chunk model (102 planes: five chunks of the
    509 inner planes, offsets up to 108 planes from a chunk's base): 9 FAST iterations (three passes) give the
    same bits as the two-level sweeps over the same iterations, on the benchmark IC (step case: D = 1e-4 for
    x < lx/2)."""
    n, niter = 511, 9
    print(f"[gpu] t3 full size {script} {n}^3", flush=True)
    monkeypatch.delenv("PTD_TZ", raising=False)
    spec = pkg.SCRIPTS[script]
    params = pkg.make_params(spec, n, n, n, arith=1)
    H0 = pkg.initial_condition(spec, n, n, n, kind="config")
    D = pkg.step_field(spec, n, n, n, kind="half_x") if spec["dcase"] == pkg.STEP else None
    out = []
    for t3 in ("1", "0"):
        monkeypatch.setenv("PTD_T3", t3)
        with pkg.Context(params) as ctx:
            ctx.set("H", H0)
            ctx.set("Hold", H0)
            if D is not None:
                ctx.set("D", D)
            ctx.iterate(niter)
            out.append((ctx.get("H"), ctx.get("qHz")))
    assert np.all(np.isfinite(out[0][0]))
    assert _relmax(out[0][0], H0) > 1e-9  # the sweeps moved the field
    assert np.array_equal(out[0][0], out[1][0]), f"H: rel {_relmax(out[0][0], out[1][0]):.2e}"
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("script", ["diff_3D_lin_multixpu", "diff_3D_linstep_multixpu"])
def test_t3_float32_default_rule_equals_two_level(script, pkg, monkeypatch):
    """float32 FAST linear / step solves take three-level passes by default from ~294^3 inner cells on: at 319^3
    (the float32 chunk model's depth) 9 iterations give the bits of the two-level sweeps."""
    n, niter = 319, 9
    print(f"[gpu] t3 float32 default rule {script} {n}^3", flush=True)
    monkeypatch.delenv("PTD_TZ", raising=False)
    spec = pkg.SCRIPTS[script]
    params = pkg.make_params(spec, n, n, n, arith=1, precision=pkg.FP32)
    H0 = pkg.initial_condition(spec, n, n, n, kind="config")
    D = pkg.step_field(spec, n, n, n, kind="half_x") if spec["dcase"] == pkg.STEP else None
    out = []
    for t3 in (None, "0"):
        if t3 is None:
            monkeypatch.delenv("PTD_T3", raising=False)
        else:
            monkeypatch.setenv("PTD_T3", t3)
        with pkg.Context(params) as ctx:
            ctx.set("H", H0)
            ctx.set("Hold", H0)
            if D is not None:
                ctx.set("D", D)
            ctx.iterate(niter)
            out.append((ctx.get("H"), ctx.get("qHx")))
    assert np.all(np.isfinite(out[0][0]))
    assert _relmax(out[0][0], H0) > 1e-6  # the sweeps moved the field
    assert np.array_equal(out[0][0], out[1][0]), f"H: rel {_relmax(out[0][0], out[1][0]):.2e}"
    assert np.array_equal(out[0][1], out[1][1])


def _mixed_steps(pkg, script, shape, steps):
    spec = pkg.SCRIPTS[script]
    params = pkg.make_params(spec, *shape, precision=pkg.MIXED, arith=1)
    H0 = pkg.initial_condition(spec, *shape, kind="config")
    D = pkg.step_field(spec, *shape, kind="half_x") if spec["dcase"] == pkg.STEP else None
    out = []
    with pkg.Context(params) as ctx:
        ctx.set("H", H0)
        ctx.set("Hold", H0)
        if D is not None:
            ctx.set("D", D)
        for _ in range(steps):
            n, rms, mx = ctx.pt_solve()
            out.append((int(n), int(ctx.outer_iterations), float(rms), float(mx)))
            ctx.assign()
        return out, ctx.get("H")


@pytest.mark.parametrize("shape,steps", [((45, 38, 29), 3), ((131, 19, 47), 2), ((319, 319, 319), 1)],
                         ids=["45x38x29", "131x19x47", "319^3"])
@pytest.mark.parametrize("script", ["diff_3D_lin_multixpu", "diff_3D_linstep_multixpu"])
def test_t3_mixed_correction_passes_equal_two_level(script, shape, steps, pkg, monkeypatch):
    """The mixed solve's float32 correction sweeps for linear / step D as three-level passes (forced on the
    small grids, the default rule at 319^3) give the float32 sweep counts, correction counts, residuals and
    final H of the two-level correction sweeps bit for bit."""
    print(f"[gpu] t3 mixed correction passes {script} {shape}", flush=True)
    if shape[0] < 300:
        monkeypatch.setenv("PTD_T3", "1")
    else:
        monkeypatch.delenv("PTD_T3", raising=False)
    got, Hg = _mixed_steps(pkg, script, shape, steps)
    monkeypatch.setenv("PTD_T3", "0")
    ref, Hr = _mixed_steps(pkg, script, shape, steps)
    assert got == ref
    assert np.array_equal(Hg, Hr)


@pytest.mark.parametrize("script", ["diff_3D_nonlin_multixpu", "diff_3D_nonlin_perf"])
def test_t3_nonlinear_model_choice_equals_two_level(script, pkg, monkeypatch):
    """float64 FAST D = H^3 takes three-level passes where the model of rounds of blocks favours them (112^3
    is such a grid): a 2-step solve gives the iteration counts, errors and bits of the two-level sweeps."""
    n, nt = 112, 2
    print(f"[gpu] t3 nonlinear model choice {script} {n}^3", flush=True)
    monkeypatch.delenv("PTD_TZ", raising=False)
    spec = pkg.SCRIPTS[script]
    params = pkg.make_params(spec, n, n, n, arith=1)
    H0 = pkg.initial_condition(spec, n, n, n, kind="config")
    out = []
    for t3 in (None, "0"):
        if t3 is None:
            monkeypatch.delenv("PTD_T3", raising=False)
        else:
            monkeypatch.setenv("PTD_T3", t3)
        with pkg.Context(params) as ctx:
            ctx.set("H", H0)
            ctx.set("Hold", H0)
            it, er = ctx.solve(nt)
            out.append((list(it), list(er), ctx.get("H")))
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])


def test_iterations_per_launch_follows_the_rules(pkg, monkeypatch):
    """ptd_ctx_iterations_per_launch reports the sweep a context's unchecked iterations run: three-level passes
    for float64 FAST linear / step D and, where the model of rounds favours them, D = H^3 (511^3, 383^3 and 319^3 yes,
    255^3 no); float32 D = H^3 and EXACT two-level; PTD_T3=0 two-level; small 2D grids LDS-resident (10)."""
    monkeypatch.delenv("PTD_T3", raising=False)
    cases = [("diff_3D_nonlin_perf", (511, 511, 511), pkg.FP64, 1, 3),
             ("diff_3D_nonlin_perf", (255, 255, 255), pkg.FP64, 1, 2),
             ("diff_3D_nonlin_perf", (319, 319, 319), pkg.FP64, 1, 3),
             ("diff_3D_nonlin_perf", (383, 383, 383), pkg.FP64, 1, 3),
             ("diff_3D_nonlin_perf", (112, 112, 112), pkg.FP64, 1, 3),
             ("diff_3D_nonlin_perf", (511, 511, 511), pkg.FP32, 1, 2),
             ("diff_3D_nonlin_perf", (511, 511, 511), pkg.FP64, 0, 2),
             ("diff_3D_lin_multixpu", (95, 95, 95), pkg.FP64, 1, 3),
             ("diff_3D_linstep_multixpu", (319, 319, 319), pkg.FP32, 1, 3),
             ("diff_2D_lin", (512, 512), pkg.FP64, 0, 10)]
    for script, shape, prec, arith, want in cases:
        params = pkg.make_params(pkg.SCRIPTS[script], *shape, precision=prec, arith=arith)
        with pkg.Context(params) as ctx:
            assert ctx.iterations_per_launch() == want, (script, shape, prec, arith)
    monkeypatch.setenv("PTD_T3", "0")
    params = pkg.make_params(pkg.SCRIPTS["diff_3D_nonlin_perf"], 511, 511, 511, arith=1)
    with pkg.Context(params) as ctx:
        assert ctx.iterations_per_launch() == 2
