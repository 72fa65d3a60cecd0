"""Pins of the oracle beyond the fixtures: what the mathematics fixes.

* converged PT solutions equal a direct solve of the backward-Euler system (tridiagonal / Newton);
* 1D linear diffusion approaches the analytic solution printed by scripts/diff_1D_lin.jl:80;
* Re = pi + sqrt(pi^2 + Da) minimises the PT iteration count (dispersion analysis,
  dispersion_analysis/dispersion_analysis_transient_diffusion1D.m);
* invariants: Dirichlet boundaries never change, x<->y symmetry, a constant state is a fixed point.
"""
import math

import numpy as np
import pytest


def _be_step_linear(Hold, Dface, dx, dt):
    """Direct solve of (H - Hold)/dt = d/dx(Dface dH/dx) on the inner cells, boundaries fixed."""
    n = Hold.size
    m = n - 2
    A = np.zeros((m, m))
    b = Hold[1:-1] / dt
    for c in range(m):
        i = c + 1
        dl, dr = Dface[i - 1], Dface[i]
        A[c, c] = 1 / dt + (dl + dr) / dx**2
        if c > 0:
            A[c, c - 1] = -dl / dx**2
        else:
            b[c] += dl / dx**2 * Hold[0]
        if c < m - 1:
            A[c, c + 1] = -dr / dx**2
        else:
            b[c] += dr / dx**2 * Hold[-1]
    H = Hold.copy()
    H[1:-1] = np.linalg.solve(A, b)
    return H


@pytest.mark.parametrize("script", ["diff_1D_lin", "diff_1D_linstep"])
def test_pt_converges_to_backward_euler_solution(script, orc):
    n = 64
    r = orc.run_script(script, n)
    case = r["case"]
    spec = orc.SCRIPTS[script]
    d = orc.derive(case)
    if spec["dcase"] == orc.STEP:
        Dc = r["D"]
        Dface = (Dc[:-1] + Dc[1:]) * 0.5
    else:
        Dface = np.full(n - 1, spec["d0"])
    H = r["H0"].copy()
    for _ in range(case.nt):
        H = _be_step_linear(H, Dface, d.dx, case.dt)
    # PT stops at RMS residual 1e-8; the solution error is of the order tol*dt
    assert np.max(np.abs(H - r["H"])) < 2e-8


def test_nonlinear_pt_converges_to_newton_solution(orc):
    n = 64
    r = orc.run_script("diff_1D_nonlin", n)
    case = r["case"]
    d = orc.derive(case)
    dx, dt = d.dx, case.dt
    H = r["H0"].copy()
    for _ in range(case.nt):
        Hold = H.copy()
        for _ in range(50):  # Newton on F(H) = (H-Hold)/dt + d/dx(q), q = -((Ha+Hb)/2)^3 dH/dx
            av = (H[:-1] + H[1:]) * 0.5
            q = -(av ** 3) * (H[1:] - H[:-1]) / dx
            F = (H[1:-1] - Hold[1:-1]) / dt + (q[1:] - q[:-1]) / dx
            # Jacobian of q_f w.r.t. H_a (left) and H_b (right)
            dq_da = -(3 * av**2 * 0.5) * (H[1:] - H[:-1]) / dx + av**3 / dx
            dq_db = -(3 * av**2 * 0.5) * (H[1:] - H[:-1]) / dx - av**3 / dx
            m = n - 2
            J = np.zeros((m, m))
            for c in range(m):
                i = c + 1
                J[c, c] = 1 / dt + (dq_da[i] - dq_db[i - 1]) / dx
                if c > 0:
                    J[c, c - 1] = -dq_da[i - 1] / dx
                if c < m - 1:
                    J[c, c + 1] = dq_db[i] / dx
            delta = np.linalg.solve(J, -F)
            H[1:-1] += delta
            if np.max(np.abs(delta)) < 1e-15:
                break
    assert np.max(np.abs(H - r["H"])) < 2e-8


def test_1d_linear_close_to_analytic(orc):
    """Hana = 1/sqrt(4(ttot+1/4)) exp(-(x-lx/2)^2/(4D(ttot+1/4))) (scripts/diff_1D_lin.jl:80); the
    discrete solution converges to it as dx, dt -> 0 (here first order in dt dominates)."""
    errs = []
    for n in (64, 256):
        r = orc.run_script("diff_1D_lin", n)
        x = r["xc"]
        ttot = 1.0
        Hana = 1 / math.sqrt(4 * (ttot + 0.25)) * np.exp(-(x - 5.0) ** 2 / (4 * (ttot + 0.25)))
        errs.append(np.linalg.norm(r["H"] - Hana) / math.sqrt(n))
    assert errs[0] < 0.05 and errs[1] < 0.05


def test_re_optimum_from_dispersion_analysis(orc):
    """The spectral radius of the 1D linear PT iteration operator is smaller at
    Re = pi + sqrt(pi^2 + Da), Da = lx^2/(D dt), than at 0.7x or 1.4x that value."""
    spec = dict(orc.SCRIPTS["diff_1D_lin"])
    n = 256
    case = orc.make_case(spec, n)
    d = orc.derive(case)
    dx, dt, vp, lx = d.dx, case.dt, d.vpdt, case.lx
    m = n - 2
    nq = m + 1
    G = np.zeros((nq, m))      # face gradient from the inner H (boundary H is fixed: affine part dropped)
    for f in range(nq):
        if f < m:
            G[f, f] += 1 / dx
        if f >= 1:
            G[f, f - 1] -= 1 / dx
    Dv = np.zeros((m, nq))     # cell divergence of the face fluxes
    for c in range(m):
        Dv[c, c + 1] += 1 / dx
        Dv[c, c] -= 1 / dx

    def rho(re):
        th = lx / vp / re
        dtau = vp * lx / 1.0 / re
        Aqq = np.eye(nq) * th / (1 + th)
        AqH = -G / (1 + th)
        s = 1 / (1 + dtau / dt)
        A = np.block([[s * (np.eye(m) - dtau * Dv @ AqH), -s * dtau * Dv @ Aqq], [AqH, Aqq]])
        return max(abs(np.linalg.eigvals(A)))

    re_opt = math.pi + math.sqrt(math.pi ** 2 + lx ** 2 / 1.0 / dt)
    assert abs(d.re - re_opt) <= 1e-14 * re_opt
    r_opt = rho(re_opt)
    assert r_opt < 1.0
    assert r_opt < rho(0.7 * re_opt) and r_opt < rho(1.4 * re_opt)


@pytest.mark.parametrize("script,n", [("diff_2D_nonlin", 48), ("diff_3D_nonlin_perf", 24), ("diff_3D_lin_multixpu", 24)])
def test_boundaries_fixed_and_xy_symmetric(script, n, orc):
    r = orc.run_script(script, n, n, n, **({"itmax": 60} if "perf" in script else {}))
    H, H0 = r["H"], r["H0"]
    nd = r["case"].ndim
    for d in range(nd):
        for edge in (0, n - 1):
            sl = [slice(None)] * nd
            sl[d] = edge
            assert np.array_equal(H[tuple(sl)], H0[tuple(sl)])
    # the IC is symmetric in x <-> y (same axis rule on both), so every PT iterate is too, bit for bit
    assert np.array_equal(H, np.swapaxes(H, 0, 1))


def test_constant_state_is_a_fixed_point(orc):
    spec = dict(orc.SCRIPTS["diff_3D_nonlin_multixpu"])
    case = orc.make_case(spec, 20, 20, 20)
    H = np.full((20, 20, 20), 1.5, order="F")
    Hold = H.copy(order="F")
    iters, errs, st = orc.solve(case, H, Hold)
    assert np.all(H == 1.5)
    assert list(iters) == [10] * case.nt and all(e == 0.0 for e in errs)
