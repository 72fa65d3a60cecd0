"""The dispersion-analysis utility (Re_opt) against what fixes it independently: the numerical optimum
of the spectral abscissa (the reference's own check, dispersion_analysis_transient_diffusion1D.m:27-31),
the double root at critical damping, the stationary limit, and the scripts' Re / theta_r / dtau_rho."""
import math
from importlib import import_module

import pytest

DAS = [0.0, 1e-3, 0.5, 10.0, 100.0, 500.0, 2.5e3]


@pytest.fixture(scope="module")
def disp(pkg):
    return import_module("ptd_amd.dispersion")


@pytest.mark.parametrize("da", DAS)
def test_closed_form_is_the_numerical_optimum(da, disp):
    num = disp.re_opt_numeric(da)
    assert abs(num - disp.re_opt(da)) <= 1e-6 * disp.re_opt(da)
    # and it is a strict optimum: 5% either side decays more slowly
    best = disp.spectral_abscissa(disp.re_opt(da), da)
    for f in (0.95, 1.05):
        assert disp.spectral_abscissa(f * disp.re_opt(da), da) < best


@pytest.mark.parametrize("da", DAS)
def test_critical_damping_of_the_slowest_mode(da, disp):
    re = disp.re_opt(da)
    a, b, c = disp.coefficients(re, da)
    assert abs(disp.discriminant(re, da)) <= 1e-12 * b * b
    r1, r2 = disp.roots(re, da)
    assert abs(r1 - r2) <= 1e-5 * abs(r1)                 # double root
    assert abs(r1.real - math.sqrt(da + math.pi ** 2)) <= 1e-6 * abs(r1)
    # every higher mode is underdamped and decays at the same rate: mode 1 sets the convergence
    assert disp.spectral_abscissa(re, da, kmax=50) == pytest.approx(disp.spectral_abscissa(re, da, 1), rel=1e-6)
    assert abs(disp.roots(re, da, k=2)[0].imag) > 0


@pytest.mark.parametrize("da", [0.5, 10.0, 500.0])
def test_two_critical_values_same_rate(da, disp):
    """sqrt(pi^2+Da) - pi is the other zero of the discriminant, with the same decay rate; the
    reference (and every script) takes the larger one."""
    lo, hi = disp.re_opt_lower(da), disp.re_opt(da)
    assert abs(disp.discriminant(lo, da)) <= 1e-12 * (lo + da / lo) ** 2
    assert disp.spectral_abscissa(lo, da) == pytest.approx(disp.spectral_abscissa(hi, da), rel=1e-6)
    assert hi > lo


def test_stationary_limit(disp):
    assert disp.re_opt(0.0) == disp.re_opt_stationary() == 2 * math.pi


@pytest.mark.parametrize("script", ["diff_1D_lin", "diff_2D_lin", "diff_3D_lin_multixpu"])
def test_scripts_use_re_opt(script, pkg, disp):
    """The scripts' Re = pi + sqrt(pi^2 + max_l^2/D/dt) is Re_opt at Da = max_l^2/(D dt), and their
    theta_r_dtau / dtau_rho follow from it (bit for bit in the scripts' operation order)."""
    spec = pkg.SCRIPTS[script]
    n = {1: (64, 1, 1), 2: (64, 64, 1), 3: (32, 32, 32)}[spec["ndim"]]
    p = pkg.make_params(spec, *n)
    ref = disp.pt_parameters(p.max_l, spec["D"], spec["dt"], p.vpdt)
    assert p.re == ref["re"] and p.theta_r == ref["theta_r"] and p.dtau_rho == ref["dtau_rho"]
    assert p.re == pytest.approx(disp.re_opt(disp.damkoehler(p.max_l, spec["D"], spec["dt"])), rel=1e-15)


def test_planted_wrong_formula_is_not_optimal(disp):
    """A plausible slip (pi + sqrt(Da), dropping pi^2) is measurably off the numerical optimum."""
    for da in (0.5, 10.0, 100.0):
        wrong = math.pi + math.sqrt(da)
        assert abs(wrong - disp.re_opt_numeric(da)) > 1e-3 * disp.re_opt(da)
