"""The reference for gradients of the lin / step solves (oracle/be_linear.py) against what fixes it
independently: its backward-Euler solution is the C oracle's converged PT solve of the reference
scripts; its adjoint gradient matches central finite differences of exact solves; its forward mode
(tangent) matches the adjoint; planted mistakes in the gradient fail the finite-difference check."""
import numpy as np
import pytest

from oracle.be_linear import BELinear


def _problem(orc, script, n):
    r = orc.run_script(script, *n)
    case = r["case"]
    d = orc.derive(case)
    h = [d.dx, d.dy, d.dz][:case.ndim]
    D = r["D"] if r["D"] is not None else np.full(r["H0"].shape, case.d0 if hasattr(case, "d0") else 1.0, order="F")
    return r, case, h, D


@pytest.mark.parametrize("script,n", [("diff_2D_linstep", (24, 20)), ("diff_3D_linstep_multixpu", (14, 12, 10)),
                                      ("diff_2D_lin", (22, 22))])
def test_backward_euler_is_the_converged_pt_solve(script, n, orc):
    r, case, h, D = _problem(orc, script, n)
    be = BELinear(D, h, case.dt)
    Hs = be.forward(r["H0"], len(r["iters"]))
    # the PT solve stops at an RMS residual below 1e-8; the error in H is of order tol * dt
    assert np.max(np.abs(Hs[-1] - r["H"])) <= 2e-8 * max(1.0, case.dt)


def _random_problem(shape, seed):
    rng = np.random.default_rng(seed)
    D = np.asfortranarray(np.exp(0.3 * rng.standard_normal(shape)))
    H0 = np.asfortranarray(rng.random(shape))
    Hobs = np.asfortranarray(rng.random(shape))
    h = [0.3, 0.25, 0.2][:len(shape)]
    return D, H0, Hobs, h, 0.2


@pytest.mark.parametrize("shape", [(9, 8), (7, 6, 5)])
def test_adjoint_gradient_matches_finite_differences(shape):
    D, H0, Hobs, h, dt = _random_problem(shape, 1)
    nt = 3
    L, gD, gH0 = BELinear(D, h, dt).gradient(H0, Hobs, nt)
    rng = np.random.default_rng(2)
    for _ in range(4):
        vD = np.asfortranarray(rng.standard_normal(shape))
        vH = np.asfortranarray(rng.standard_normal(shape))
        eps = 1e-6
        Lp = BELinear(D + eps * vD, h, dt).forward(H0 + eps * vH, nt)[-1]
        Lm = BELinear(D - eps * vD, h, dt).forward(H0 - eps * vH, nt)[-1]
        be = BELinear(D, h, dt)
        fd = (be.loss(Lp, Hobs) - be.loss(Lm, Hobs)) / (2 * eps)
        ad = float(np.sum(gD * vD) + np.sum(gH0 * vH))
        assert abs(fd - ad) <= 1e-7 * max(1.0, abs(ad)), (fd, ad)


@pytest.mark.parametrize("shape", [(9, 8), (7, 6, 5)])
def test_forward_mode_matches_adjoint(shape):
    D, H0, Hobs, h, dt = _random_problem(shape, 3)
    be = BELinear(D, h, dt)
    L, gD, gH0 = be.gradient(H0, Hobs, 4)
    rng = np.random.default_rng(4)
    for _ in range(5):
        vD = rng.standard_normal(shape)
        vH = rng.standard_normal(shape)
        jv = be.jvp(H0, Hobs, 4, vD, vH)
        assert abs(jv - (np.sum(gD * vD) + np.sum(gH0 * vH))) <= 1e-11 * max(1.0, abs(jv))


def test_planted_gradient_mistakes_fail_finite_differences():
    D, H0, Hobs, h, dt = _random_problem((8, 7), 5)
    be = BELinear(D, h, dt)
    L, gD, gH0 = be.gradient(H0, Hobs, 2)
    rng = np.random.default_rng(6)
    vD = rng.standard_normal(D.shape)
    eps = 1e-6
    fd = (be.loss(BELinear(D + eps * vD, h, dt).forward(H0, 2)[-1], Hobs)
          - be.loss(BELinear(D - eps * vD, h, dt).forward(H0, 2)[-1], Hobs)) / (2 * eps)
    good = float(np.sum(gD * vD))
    assert abs(fd - good) <= 1e-7 * abs(good)
    for wrong in (2.0 * gD, -gD):                 # the face sensitivity not halved / the wrong sign
        assert abs(fd - float(np.sum(wrong * vD))) > 1e-3 * abs(good)


# ---- the nonlinear case (oracle/be_nonlin.py) ------------------------------------------------------

from oracle.be_nonlin import BENonlin  # noqa: E402


@pytest.mark.parametrize("script,n", [("diff_1D_nonlin", (40,)), ("diff_2D_nonlin", (24, 20)),
                                      ("diff_3D_nonlin_multixpu", (14, 12, 10))])
def test_nonlin_backward_euler_is_the_converged_pt_solve(script, n, orc):
    """Newton's backward-Euler steps of F(H) = (H - Hold)/dt + div q(H) = 0 reproduce the C oracle's converged
    PT solve of the reference's nonlinear scripts (its RMS residual < 1e-8 leaves an error of order tol*dt)."""
    r = orc.run_script(script, *n)
    case = r["case"]
    d = orc.derive(case)
    h = [d.dx, d.dy, d.dz][:case.ndim]
    be = BENonlin(r["H0"].shape, h, case.dt)
    Hs = be.forward(r["H0"], len(r["iters"]))
    assert np.max(np.abs(Hs[-1] - r["H"])) <= 2e-8 * max(1.0, case.dt)
    assert np.max(np.abs(Hs[-1] - r["H0"])) > 1e-4  # the steps moved the field


def _nonlin_problem(shape, seed):
    rng = np.random.default_rng(seed)
    H0 = np.asfortranarray(1.0 + 0.3 * rng.random(shape))
    Hobs = np.asfortranarray(1.0 + 0.3 * rng.random(shape))
    h = [0.3, 0.25, 0.2][:len(shape)]
    return H0, Hobs, h, 0.2


@pytest.mark.parametrize("shape", [(12,), (9, 8), (7, 6, 5)])
def test_nonlin_adjoint_gradient_matches_finite_differences(shape):
    H0, Hobs, h, dt = _nonlin_problem(shape, 4)
    nt = 3
    be = BENonlin(shape, h, dt)
    L, g = be.gradient(H0, Hobs, nt)
    assert abs(L - be.loss(be.forward(H0, nt)[-1], Hobs)) <= 1e-12 * L
    rng = np.random.default_rng(5)
    for _ in range(4):
        v = np.asfortranarray(rng.standard_normal(shape))
        eps = 1e-6
        fd = (be.loss(be.forward(H0 + eps * v, nt)[-1], Hobs) - be.loss(be.forward(H0 - eps * v, nt)[-1], Hobs)) / (2 * eps)
        ad = float(np.sum(g * v))
        assert abs(fd - ad) <= 1e-7 * max(1.0, abs(ad)), (fd, ad)


@pytest.mark.parametrize("shape", [(9, 8), (7, 6, 5)])
def test_nonlin_forward_mode_matches_adjoint(shape):
    H0, Hobs, h, dt = _nonlin_problem(shape, 6)
    be = BENonlin(shape, h, dt)
    _, g = be.gradient(H0, Hobs, 3)
    rng = np.random.default_rng(7)
    for _ in range(3):
        v = np.asfortranarray(rng.standard_normal(shape))
        jv = be.jvp(H0, Hobs, 3, v)
        assert abs(jv - float(np.sum(g * v))) <= 1e-11 * max(1.0, abs(jv))


def test_nonlin_planted_mistake_fails_finite_differences():
    """Dropping the dD/dH term of the Jacobian (the part that makes it non-symmetric) must be caught."""
    shape = (9, 8)
    H0, Hobs, h, dt = _nonlin_problem(shape, 4)
    nt = 3
    good = BENonlin(shape, h, dt)
    _, gbad = BENonlin(shape, h, dt, cubic_deriv=False).gradient(H0, Hobs, nt)
    v = np.asfortranarray(np.random.default_rng(8).standard_normal(shape))
    eps = 1e-6
    fd = (good.loss(good.forward(H0 + eps * v, nt)[-1], Hobs) - good.loss(good.forward(H0 - eps * v, nt)[-1], Hobs)) / (2 * eps)
    assert abs(fd - float(np.sum(gbad * v))) > 1e-3 * abs(fd)
