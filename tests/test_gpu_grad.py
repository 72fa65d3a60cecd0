"""Gradients of the lin / step solves through the C ABI (ptd_grad_solve) against the oracle's exact
adjoint and forward mode (oracle/be_linear.py, pinned by its own finite-difference tests), and against
central finite differences of GPU solves."""
import numpy as np
import pytest

from oracle.be_linear import BELinear

pytestmark = pytest.mark.gpu


def _setup(pkg, script, shape, tol, seed=0, arith=1):
    spec = pkg.SCRIPTS[script]
    p = pkg.make_params(spec, *shape, arith=arith)
    p.tol = tol
    nd = spec["ndim"]
    rng = np.random.default_rng(seed)
    ls = [spec["lx"], spec.get("ly", 0.0), spec.get("lz", 0.0)][:nd]
    r2 = 0.0
    for d in range(nd):
        dx = ls[d] / shape[d]
        xc = dx / 2 + dx * np.arange(shape[d]) - 0.5 * ls[d]
        sh = [1] * nd
        sh[d] = shape[d]
        r2 = r2 + (xc ** 2).reshape(sh)
    H0 = np.asfortranarray(np.exp(-r2 / 4.0))
    D = pkg.step_field(spec, *shape) if spec["dcase"] == pkg.STEP else np.full(shape, p.d0, order="F")
    Hobs = np.asfortranarray(H0 * (1.0 + 0.05 * rng.standard_normal(shape)))
    h = [p.dx, p.dy, p.dz][:nd]
    return p, H0, D, Hobs, h


def _grad_gpu(pkg, p, H0, D, Hobs, nt):
    with pkg.Context(p) as ctx:
        ctx.set("H", H0)
        ctx.set("Hold", H0)
        if p.dcase == pkg.STEP:
            ctx.set("D", D)
        return ctx.grad_solve(nt, Hobs)


@pytest.mark.parametrize("script,shape", [("diff_2D_linstep", (40, 36)), ("diff_3D_linstep_multixpu", (32, 32, 32)),
                                          ("diff_2D_lin", (30, 30))])
def test_gradient_matches_oracle_adjoint_and_forward_mode(script, shape, pkg):
    print(f"[gpu] gradient vs oracle {script} {shape}", flush=True)
    nt = 4
    p, H0, D, Hobs, h = _setup(pkg, script, shape, tol=1e-12)
    g = _grad_gpu(pkg, p, H0, D, Hobs, nt)
    be = BELinear(D, h, p.dt)
    L, gD, gH0 = be.gradient(H0, Hobs, nt)
    assert abs(g["loss"] - L) <= 1e-9 * L
    assert np.max(np.abs(g["grad_D"] - gD)) <= 1e-6 * np.max(np.abs(gD))
    assert np.max(np.abs(g["grad_H0"] - gH0)) <= 1e-6 * np.max(np.abs(gH0))
    rng = np.random.default_rng(7)
    for _ in range(20):  # forward mode of the reference system along random directions
        vD = rng.standard_normal(shape)
        vH = rng.standard_normal(shape)
        jv = be.jvp(H0, Hobs, nt, vD, vH)
        gv = float(np.sum(g["grad_D"] * vD) + np.sum(g["grad_H0"] * vH))
        assert abs(gv - jv) <= 1e-6 * max(abs(jv), 1e-30), (gv, jv)


def test_gradient_matches_finite_differences_on_gpu(pkg):
    print("[gpu] gradient vs finite differences 48^3", flush=True)
    script, shape, nt = "diff_3D_linstep_multixpu", (48, 48, 48), 3
    p, H0, D, Hobs, h = _setup(pkg, script, shape, tol=1e-12, seed=1)
    g = _grad_gpu(pkg, p, H0, D, Hobs, nt)
    inner = (slice(1, -1),) * 3

    def loss(Dv, H0v):
        with pkg.Context(p) as ctx:
            ctx.set("H", H0v)
            ctx.set("Hold", H0v)
            ctx.set("D", Dv)
            ctx.solve(nt)
            r = ctx.get("H")[inner] - Hobs[inner]
            return 0.5 * float(np.sum(r * r))

    rng = np.random.default_rng(11)
    for _ in range(4):
        vD = np.asfortranarray(rng.standard_normal(shape) * D)   # relative perturbations of D
        vH = np.asfortranarray(rng.standard_normal(shape) * 0.1)
        eps = 1e-5
        fd = (loss(D + eps * vD, H0 + eps * vH) - loss(D - eps * vD, H0 - eps * vH)) / (2 * eps)
        gv = float(np.sum(g["grad_D"] * vD) + np.sum(g["grad_H0"] * vH))
        assert abs(fd - gv) <= 1e-5 * abs(gv), (fd, gv)


@pytest.mark.parametrize("script,shape", [("diff_3D_linstep_multixpu", (32, 32, 32)), ("diff_2D_lin", (40, 36))])
def test_gradient_of_a_small_misfit_at_the_default_tol(script, shape, pkg):
    """H_obs = H(T)*(1 + 1e-6) with the scripts' tol 1e-8: the adjoint right-hand side (H(T) - H_obs)*dt is
    far below tol in absolute terms, so an adjoint stopped on the forward's absolute tol would end at its
    first check. The adjoint solves stop relative to their right-hand side instead: the gradient matches
    the exact adjoint of the same misfit (the reference's H_obs shifted by the exact and GPU H(T)'s
    difference, so that both sides differentiate the same residual) to 1e-4."""
    print(f"[gpu] gradient of a 1e-6 misfit {script} {shape}", flush=True)
    nt = 3
    p, H0, D, _, h = _setup(pkg, script, shape, tol=1e-8)
    with pkg.Context(p) as ctx:
        ctx.set("H", H0)
        ctx.set("Hold", H0)
        if p.dcase == pkg.STEP:
            ctx.set("D", D)
        ctx.solve(nt)
        Hg = ctx.get("H")
    Hobs = np.asfortranarray(Hg * (1.0 + 1e-6))
    g = _grad_gpu(pkg, p, H0, D, Hobs, nt)
    assert np.all(g["iters_adj"] > 10)
    be = BELinear(D, h, p.dt)
    Hbe = be.forward(H0, nt)[-1]
    L, gD, gH0 = be.gradient(H0, np.asfortranarray(Hobs + (Hbe - Hg)), nt)
    assert abs(g["loss"] - L) <= 1e-9 * L
    assert np.max(np.abs(g["grad_D"] - gD)) <= 1e-4 * np.max(np.abs(gD))
    assert np.max(np.abs(g["grad_H0"] - gH0)) <= 1e-4 * np.max(np.abs(gH0))


def test_context_ends_as_after_solve(pkg):
    script, shape, nt = "diff_3D_linstep_multixpu", (30, 26, 22), 3
    p, H0, D, Hobs, h = _setup(pkg, script, shape, tol=1e-8)
    states = []
    for use_grad in (True, False):
        with pkg.Context(p) as ctx:
            ctx.set("H", H0)
            ctx.set("Hold", H0)
            ctx.set("D", D)
            its = ctx.grad_solve(nt, Hobs)["iters_fwd"] if use_grad else ctx.solve(nt)[0]
            states.append((its, ctx.get("H"), ctx.get("Hold"), ctx.get("qHx"), ctx.get("qHz"), ctx.solve(1)[0]))
    for a, b in zip(*states):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("script,shape", [("diff_3D_linstep_multixpu", (3, 3, 3)), ("diff_3D_lin_multixpu", (5, 6, 3)),
                                          ("diff_2D_linstep", (4, 3))])
def test_gradient_on_the_smallest_grids(script, shape, pkg):
    """One inner cell / one inner row: the adjoint source, the face gather into the cells and the boundary
    coupling of dL/dH0 on grids whose every face touches the boundary."""
    print(f"[gpu] gradient on {script} {shape}", flush=True)
    nt = 3
    p, H0, D, Hobs, h = _setup(pkg, script, shape, tol=1e-12)
    g = _grad_gpu(pkg, p, H0, D, Hobs, nt)
    be = BELinear(D, h, p.dt)
    L, gD, gH0 = be.gradient(H0, Hobs, nt)
    assert abs(g["loss"] - L) <= 1e-9 * max(L, 1e-300)
    assert np.max(np.abs(g["grad_D"] - gD)) <= 1e-6 * max(np.max(np.abs(gD)), 1e-300)
    assert np.max(np.abs(g["grad_H0"] - gH0)) <= 1e-6 * max(np.max(np.abs(gH0)), 1e-300)


# ---- the nonlinear case (D = H^3): the transposed-Jacobian adjoint ----------------------------------

from oracle.be_nonlin import BENonlin  # noqa: E402


def _setup_nonlin(pkg, script, shape, tol, seed=0):
    """H0 = 0.5 + exp(-r^2/4) (positive everywhere, so D = H^3 stays away from 0), H_obs a 5% perturbation."""
    spec = pkg.SCRIPTS[script]
    p = pkg.make_params(spec, *shape, arith=1)
    p.tol = tol
    nd = spec["ndim"]
    ls = [spec["lx"], spec.get("ly", 0.0), spec.get("lz", 0.0)][:nd]
    r2 = 0.0
    for d in range(nd):
        dx = ls[d] / shape[d]
        xc = dx / 2 + dx * np.arange(shape[d]) - 0.5 * ls[d]
        sh = [1] * nd
        sh[d] = shape[d]
        r2 = r2 + (xc ** 2).reshape(sh)
    H0 = np.asfortranarray(0.5 + np.exp(-r2 / 4.0))
    rng = np.random.default_rng(seed)
    Hobs = np.asfortranarray(H0 * (1.0 + 0.05 * rng.standard_normal(shape)))
    h = [p.dx, p.dy, p.dz][:nd]
    return p, H0, Hobs, h


@pytest.mark.parametrize("script,shape", [("diff_1D_nonlin", (64,)), ("diff_2D_nonlin", (40, 36)),
                                          ("diff_3D_nonlin_multixpu", (24, 22, 20))])
def test_nonlin_gradient_matches_oracle_adjoint_and_forward_mode(script, shape, pkg):
    """dL/dH0 through the transposed-Jacobian adjoint against the exact discrete adjoint of Newton's
    backward-Euler steps (oracle/be_nonlin.py) and 10 forward-mode directional derivatives."""
    print(f"[gpu] nonlinear gradient vs oracle {script} {shape}", flush=True)
    nt = 3
    p, H0, Hobs, h = _setup_nonlin(pkg, script, shape, tol=1e-12)
    g = _grad_gpu(pkg, p, H0, None, Hobs, nt)
    assert g["grad_D"] is None
    be = BENonlin(shape, h, p.dt)
    L, gH0 = be.gradient(H0, Hobs, nt)
    assert abs(g["loss"] - L) <= 1e-9 * L
    assert np.max(np.abs(g["grad_H0"] - gH0)) <= 1e-6 * np.max(np.abs(gH0)), \
        np.max(np.abs(g["grad_H0"] - gH0)) / np.max(np.abs(gH0))
    rng = np.random.default_rng(7)
    for _ in range(10):
        v = rng.standard_normal(shape)
        jv = be.jvp(H0, Hobs, nt, v)
        gv = float(np.sum(g["grad_H0"] * v))
        assert abs(gv - jv) <= 1e-6 * max(abs(jv), 1e-30), (gv, jv)


def test_nonlin_gradient_matches_finite_differences_on_gpu(pkg):
    print("[gpu] nonlinear gradient vs finite differences 48^3", flush=True)
    script, shape, nt = "diff_3D_nonlin_multixpu", (48, 48, 48), 3
    p, H0, Hobs, h = _setup_nonlin(pkg, script, shape, tol=1e-12, seed=1)
    g = _grad_gpu(pkg, p, H0, None, Hobs, nt)
    inner = (slice(1, -1),) * 3

    def loss(H0v):
        with pkg.Context(p) as ctx:
            ctx.set("H", H0v)
            ctx.set("Hold", H0v)
            ctx.solve(nt)
            r = ctx.get("H")[inner] - Hobs[inner]
            return 0.5 * float(np.sum(r * r))

    rng = np.random.default_rng(11)
    for _ in range(4):
        v = np.asfortranarray(rng.standard_normal(shape) * 0.1)
        eps = 1e-5
        fd = (loss(H0 + eps * v) - loss(H0 - eps * v)) / (2 * eps)
        gv = float(np.sum(g["grad_H0"] * v))
        assert abs(fd - gv) <= 1e-5 * abs(gv), (fd, gv)


def test_nonlin_gradient_context_and_arguments(pkg):
    """The context ends as after ptd_solve(nt); dL/dD is refused for D = H^3 (PTD_ERR_ARG)."""
    import ctypes
    script, shape, nt = "diff_3D_nonlin_multixpu", (30, 26, 22), 3
    p, H0, Hobs, h = _setup_nonlin(pkg, script, shape, tol=1e-8)
    states = []
    for use_grad in (True, False):
        with pkg.Context(p) as ctx:
            ctx.set("H", H0)
            ctx.set("Hold", H0)
            its = ctx.grad_solve(nt, Hobs)["iters_fwd"] if use_grad else ctx.solve(nt)[0]
            states.append((its, ctx.get("H"), ctx.get("Hold"), ctx.get("qHx"), ctx.get("qHz"), ctx.solve(1)[0]))
    for a, b in zip(*states):
        assert np.array_equal(a, b)
    with pkg.Context(p) as ctx:
        lib = pkg.load()
        gD = np.zeros(shape, order="F")
        dp = ctypes.POINTER(ctypes.c_double)
        st = lib.ptd_grad_solve(ctx.handle, nt, Hobs.ctypes.data_as(dp), Hobs.size, None, gD.ctypes.data_as(dp),
                                None, None, None)
        assert st == 1 and b"grad_d = NULL" in lib.ptd_last_error()
