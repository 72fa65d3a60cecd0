"""Dispersion analysis of the accelerated PT iteration: the optimal Reynolds number Re_opt for a given
Damkoehler number Da, restated from dispersion_analysis/dispersion_analysis_transient_diffusion1D.m
(and the stationary variant dispersion_analysis_stationary_diffusion1D.m).

The PT system (1D, pseudo-time tau)

    rho   dH/dtau + dq/dx + rho_ph H/dt = 0           (mass balance)
    theta_r dq/dtau + q + D dH/dx      = 0           (damped flux)

with V_p = sqrt(D/(rho theta_r)), Re = rho V_p L/D and Da = L^2 rho_ph/(D dt), admits Fourier modes
H ~ exp(-lambda V_p tau/L) sin(pi k x/L). Substituting them gives the dispersion relation (the quadratic
the reference's scripts reduce to with `coeffs(disp_rel, lambda_k)`):

    lambda^2 - (Re + Da/Re) lambda + (Da + pi^2 k^2) = 0.

The slowest mode (k = 1) decays fastest when it is critically damped (zero discriminant). That happens
at Re = sqrt(pi^2 + Da) -+ pi; both give the same decay rate sqrt(pi^2 + Da) (their product is Da, so
Re + Da/Re is the same), and the analysis takes the larger root (`solve(disc, Re)[2]` in the notebook):

    Re_opt = pi + sqrt(pi^2 + Da)      (Da = 0, stationary: Re_opt = 2 pi),

the value every script uses: Re = pi + sqrt(pi^2 + max_l^2/D/dt), e.g.
scripts/diff_3D/diff_3D_lin_multixpu.jl:58-62 (Da = max_l^2/(D dt) with rho_ph = 1), and for the
nonlinear case per face / cell with D = H^3. Then theta_r_dtau = max_l/Vpdtau/Re and
dtau_rho = Vpdtau*max_l/D/Re.

Host-side utility: used to choose PT parameters for new Da / physics and by the tests that pin the
scripts' Re expression to the analysis.
"""
import math

import numpy as np


def coefficients(re, da, k=1):
    """(a, b, c) of a*lambda^2 + b*lambda + c for mode k."""
    return 1.0, -(re + da / re), da + (math.pi * k) ** 2


def roots(re, da, k=1):
    """The two decay rates lambda of mode k (complex when the mode is underdamped)."""
    a, b, c = coefficients(re, da, k)
    disc = complex(b * b - 4.0 * a * c)
    s = disc ** 0.5
    return (-b - s) / (2 * a), (-b + s) / (2 * a)


def discriminant(re, da, k=1):
    a, b, c = coefficients(re, da, k)
    return b * b - 4.0 * a * c


def spectral_abscissa(re, da, kmax=1):
    """min over modes k = 1..kmax of min Re(lambda): the decay rate of the slowest component (the
    quantity the reference plots, min{Re(lambda_k)} at k = 1)."""
    return min(min(r.real for r in roots(re, da, k)) for k in range(1, kmax + 1))


def re_opt(da):
    """Closed form Re_opt = pi + sqrt(pi^2 + Da)."""
    return math.pi + math.sqrt(math.pi * math.pi + da)


def re_opt_stationary():
    """Da = 0 (dispersion_analysis_stationary_diffusion1D.m): Re_opt = 2 pi."""
    return 2.0 * math.pi


def re_opt_lower(da):
    """The other critical value, sqrt(pi^2 + Da) - pi: the same decay rate, not the one the scripts use."""
    return math.sqrt(math.pi * math.pi + da) - math.pi


def re_opt_numeric(da, lo=None, hi=None, iters=200):
    """Re maximising the spectral abscissa of mode 1 on the branch Re >= sqrt(Da), by golden-section
    search (the reference's numerical evaluation over a grid of Re, done as a 1D maximisation). On
    that branch the abscissa rises as (Re + Da/Re)/2 while the roots are complex and falls once they
    are real, so it is unimodal; the mirror branch holds the lower critical value."""
    lo = max(math.pi / 2, math.sqrt(da)) if lo is None else lo
    hi = 5 * math.pi + 2 * math.sqrt(da) if hi is None else hi
    g = (math.sqrt(5) - 1) / 2
    a, b = lo, hi
    c, d = b - g * (b - a), a + g * (b - a)
    fc, fd = spectral_abscissa(c, da), spectral_abscissa(d, da)
    for _ in range(iters):
        if fc > fd:
            b, d, fd = d, c, fc
            c = b - g * (b - a)
            fc = spectral_abscissa(c, da)
        else:
            a, c, fc = c, d, fd
            d = a + g * (b - a)
            fd = spectral_abscissa(d, da)
    return 0.5 * (a + b)


def damkoehler(max_l, d, dt, rho_ph=1.0):
    """Da = L^2 rho_ph/(D dt) as the scripts form it (max_l^2/D/dt)."""
    return max_l * max_l / d / dt * rho_ph


def pt_parameters(max_l, d, dt, vpdt):
    """Re, theta_r_dtau and dtau_rho of the linear scripts, in their operation order
    (scripts/diff_3D/diff_3D_lin_multixpu.jl:58-62)."""
    re = math.pi + math.sqrt(math.pi * math.pi + max_l * max_l / d / dt)
    return {"re": re, "theta_r": max_l / vpdt / re, "dtau_rho": vpdt * max_l / d / re}


def abscissa_map(re_grid, da_grid):
    """The reference's contour data: min Re(lambda_1) over a (Re, Da) grid."""
    out = np.empty((len(re_grid), len(da_grid)))
    for i, re in enumerate(re_grid):
        for j, da in enumerate(da_grid):
            out[i, j] = spectral_abscissa(re, da)
    return out
