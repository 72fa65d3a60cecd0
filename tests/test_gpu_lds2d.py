"""GPU parity of the LDS-resident 2D iterations (k_lds2d: up to 10 PT iterations per launch on a 32 x 32
core of cells with a 10-cell frame), against the CPU oracle and against the single sweeps.

 - EXACT arithmetic: H and both fluxes bit-identical to the oracle's niter iterations (niter = 2, 10 and
   23 = 10 + 10 + 3 per launch) on grids smaller than one core, equal to it and ragged against it, for the
   four 2D scripts (scripts/diff_2D_*.jl compute_flux! + compute_update!);
 - every precision and arithmetic: the same bits as niter single sweeps (each face and cell evaluates the
   same expressions on the same operands);
 - converged solves with checks (plain form: the checking iteration is the last of an LDS launch, whose
   per-core residual partials a one-block reduction sums; perf form: a single checking sweep after the
   LDS iterations): the
   reference's 2D linear 512^2 case to tol 1e-8 gives the oracle's 2660 iterations and its H bit for bit,
   solves of every 2D script give the same counts, bits and residual norms (to summation order) with and
   without the LDS iterations, also when a check block spans 1, 7, 10 or 13 iterations.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRIPTS_2D = ["diff_2D_lin", "diff_2D_linstep", "diff_2D_nonlin", "diff_2D_nonlin_perf"]
# one core and less, exactly one, ragged (partial cores on both axes), thin, and 4 x 2 cores
SHAPES = [(3, 3), (5, 70), (32, 32), (33, 31), (97, 130), (200, 37), (128, 64)]


def _relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _device(pkg):
    lib = pkg.load()
    if lib.ptd_device_check(0, None, 0) != 0:
        pytest.fail(f"no gfx950 device: {lib.ptd_last_error()}")


def _random_state(pkg, orc, script, shape, seed, precision=0, arith=0):
    rng = np.random.default_rng(seed)
    spec = pkg.SCRIPTS[script]
    params = pkg.make_params(spec, *shape, precision=precision, arith=arith)
    case = orc.make_case(orc.SCRIPTS[script], *shape)
    H = np.asfortranarray(rng.uniform(1.0, 1.2, size=shape))
    Hold = np.asfortranarray(rng.uniform(1.0, 1.2, size=shape))
    D = np.asfortranarray(np.where(rng.uniform(size=shape) < 0.5, 1e-4, 1.0)) if spec["dcase"] == pkg.STEP else None
    q = {name: np.asfortranarray(0.05 * rng.normal(size=pkg.field_shape(params, fid)))
         for name, fid in (("qHx", 3), ("qHy", 4))}
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
                           orc._p(qr["qHx"]), orc._p(qr["qHy"]), None, niter)
    return Hr, qr


@pytest.mark.parametrize("niter", [2, 10, 23])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("script", SCRIPTS_2D)
def test_lds_sweeps_match_oracle(script, shape, niter, pkg, orc, monkeypatch):
    print(f"[gpu] lds 2D {script} {shape} niter={niter}", flush=True)
    monkeypatch.setenv("PTD_LDS2D", "1")
    params, case, H, Hold, D, q = _random_state(pkg, orc, script, shape, seed=sum(shape) + niter)
    Hg, qg = _gpu_iterate(pkg, params, H, Hold, D, q, niter)
    Hr, qr = _oracle_iterate(orc, pkg, params, case, H, Hold, D, q, niter)
    assert np.all(np.isfinite(Hr))
    assert np.array_equal(Hg, Hr), f"H: rel {_relmax(Hg, Hr):.2e}"
    for k in q:
        assert np.array_equal(qg[k], qr[k]), f"{k}: rel {_relmax(qg[k], qr[k]):.2e}"


@pytest.mark.parametrize("precision", [0, 1], ids=["f64", "f32"])
@pytest.mark.parametrize("arith", [0, 1], ids=["exact", "fast"])
@pytest.mark.parametrize("script", SCRIPTS_2D)
def test_lds_equals_single_sweeps(script, arith, precision, pkg, orc, monkeypatch):
    shape, niter = (97, 130), 23
    print(f"[gpu] lds 2D vs single sweeps {script} arith={arith} precision={precision}", flush=True)
    params, case, H, Hold, D, q = _random_state(pkg, orc, script, shape, seed=5, precision=precision, arith=arith)
    out = []
    for lds in ("1", "0"):
        monkeypatch.setenv("PTD_LDS2D", lds)
        monkeypatch.setenv("PTD_T2", "0")
        out.append(_gpu_iterate(pkg, params, H, Hold, D, q, niter))
    assert np.all(np.isfinite(out[1][0]))
    assert np.array_equal(out[0][0], out[1][0]), f"H: rel {_relmax(out[0][0], out[1][0]):.2e}"
    for k in q:
        assert np.array_equal(out[0][1][k], out[1][1][k]), k


def test_lds_config_512_matches_oracle(pkg, orc):
    """The reference's CPU-runnable case (2D linear, 512^2, 5 steps to tol 1e-8, checks every 10
    iterations): 2660 iterations and the oracle's H, bit for bit."""
    print("[gpu] lds 2D config 512^2 vs oracle", flush=True)
    ref = orc.run_script("diff_2D_lin", 512, 512)
    res = pkg.run_script("diff_2D_lin", 512, 512, H0=ref["H0"])
    assert int(np.sum(ref["iters"])) == 2660
    assert list(res.iters) == list(ref["iters"])
    assert np.array_equal(res.H, ref["H"]), f"rel {_relmax(res.H, ref['H']):.2e}"


@pytest.mark.parametrize("arith", [0, 1], ids=["exact", "fast"])
@pytest.mark.parametrize("script", SCRIPTS_2D)
def test_lds_solve_equals_single_sweeps(script, arith, pkg, monkeypatch):
    """Converged 5-step solves at 150 x 140: the same per-step counts and H with and without the LDS
    iterations; residual norms agree to summation order."""
    shape = (150, 140)
    print(f"[gpu] lds 2D solve {script} arith={arith}", flush=True)
    spec = pkg.SCRIPTS[script]
    over = dict(nout=10, itMax=1e5) if "perf" in script else {}
    out = []
    for lds in ("1", "0"):
        monkeypatch.setenv("PTD_LDS2D", lds)
        H0 = pkg.initial_condition(spec, *shape)
        D = pkg.step_field(spec, *shape) if spec["dcase"] == pkg.STEP else None
        out.append(pkg.run_script(script, *shape, H0=H0, D=D, arith=arith, **over))
    assert list(out[0].iters) == list(out[1].iters) and int(np.sum(out[0].iters)) > 0
    assert np.array_equal(out[0].H, out[1].H), f"rel {_relmax(out[0].H, out[1].H):.2e}"
    np.testing.assert_allclose(out[0].errs, out[1].errs, rtol=1e-12)


@pytest.mark.parametrize("nout", [1, 7, 10, 13])
@pytest.mark.parametrize("script", ["diff_2D_lin", "diff_2D_nonlin", "diff_2D_linstep"])
def test_lds_checks_every_nout(script, nout, pkg, monkeypatch):
    """Check blocks of 1, 7, 10 and 13 iterations (13 = one LDS launch of 10 and a checking launch of 3):
    counts, H and residual norms as with single sweeps."""
    shape = (150, 140)
    print(f"[gpu] lds 2D checks {script} nout={nout}", flush=True)
    spec = pkg.SCRIPTS[script]
    out = []
    for lds in ("1", "0"):
        monkeypatch.setenv("PTD_LDS2D", lds)
        H0 = pkg.initial_condition(spec, *shape)
        D = pkg.step_field(spec, *shape) if spec["dcase"] == pkg.STEP else None
        out.append(pkg.run_script(script, *shape, H0=H0, D=D, nout=nout))
    assert list(out[0].iters) == list(out[1].iters) and int(np.sum(out[0].iters)) > 0
    assert all(int(v) % nout == 0 for v in out[0].iters)
    assert np.array_equal(out[0].H, out[1].H), f"rel {_relmax(out[0].H, out[1].H):.2e}"
    np.testing.assert_allclose(out[0].errs, out[1].errs, rtol=1e-12)


@pytest.mark.parametrize("itmax", [30, 34, 100000])
@pytest.mark.parametrize("script", ["diff_2D_lin", "diff_2D_nonlin"])
def test_pipelined_checks_equal_blocking_loop(script, itmax, pkg, monkeypatch):
    """Small 2D solves enqueue the next check block before reading the last residual: counts, residual
    norms and H equal those of the loop that waits at every check, also when itMax (a multiple of nout or
    not) ends the solve before convergence."""
    shape = (120, 100)
    print(f"[gpu] pipelined checks {script} itmax={itmax}", flush=True)
    spec = pkg.SCRIPTS[script]
    out = []
    for pipe in ("1", "0"):
        monkeypatch.setenv("PTD_PIPELINE", pipe)
        H0 = pkg.initial_condition(spec, *shape)
        out.append(pkg.run_script(script, *shape, H0=H0, itMax=itmax))
    assert list(out[0].iters) == list(out[1].iters)
    if itmax < 1000:
        assert all(int(v) == itmax for v in out[0].iters)
    assert np.array_equal(out[0].errs, out[1].errs)
    assert np.array_equal(out[0].H, out[1].H)
