"""Host-side logic of the package (parameters, initial conditions, D fields, shapes) against the oracle,
on the CPU."""
import re

import numpy as np
import pytest

from conftest import FIXTURE_SCRIPTS, ROOT


def _script_dims(name):
    return {1: (64, 1, 1), 2: (64, 64, 1), 3: (32, 32, 32)}[int(name.split("_")[1][0])]


@pytest.mark.parametrize("script", sorted({s for s, _ in FIXTURE_SCRIPTS.values()}))
def test_derived_numerics_match_oracle(script, pkg, orc):
    nx, ny, nz = _script_dims(script)
    p = pkg.make_params(pkg.SCRIPTS[script], nx, ny, nz)
    case = orc.make_case(orc.SCRIPTS[script], nx, ny, nz)
    d = orc.derive(case)
    assert p.dx == d.dx and p.vpdt == d.vpdt and p.max_l == d.max_l and p.max_l2 == d.max_l2
    assert p.dt == case.dt and p.resc == case.resc and p.nout == case.nout and p.itmax == case.itmax
    if p.dcase != pkg.NONLIN:
        assert p.re == d.re and p.theta_r == d.theta_r
    if p.dcase == pkg.LIN:
        assert p.dtau_rho == d.dtau_rho
    if p.ndim >= 2:
        assert p.dy == d.dy
    if p.ndim >= 3:
        assert p.dz == d.dz


def test_linrange_matches_oracle(pkg, orc):
    from importlib import import_module
    scripts = import_module("ptd_amd.scripts")
    for n in (3, 32, 64, 511, 1023):
        a, b = 10.0 / n / 2, 10.0 - 10.0 / n / 2
        assert np.array_equal(scripts.linrange(a, b, n), orc.linrange(a, b, n))


@pytest.mark.parametrize("script", sorted({s for s, _ in FIXTURE_SCRIPTS.values()}))
def test_initial_condition_matches_oracle(script, pkg, orc):
    nx, ny, nz = _script_dims(script)
    spec = pkg.SCRIPTS[script]
    H0 = pkg.initial_condition(spec, nx, ny, nz)
    case = orc.make_case(orc.SCRIPTS[script], nx, ny, nz)
    ref = orc.initial_condition(orc.SCRIPTS[script]["ic"], case)
    assert H0.shape == ref.shape and H0.flags["F_CONTIGUOUS"]
    # numpy's exp may differ from libm's by one ulp; everything before exp is bit-identical
    ulp = np.spacing(np.abs(ref))
    assert np.all(np.abs(H0 - ref) <= ulp)


def test_config_initial_condition(pkg, orc):
    spec = pkg.SCRIPTS["diff_3D_nonlin_perf"]
    H0 = pkg.initial_condition(spec, 21, 21, 21, kind="config", amp=2.0, bval=1.0)
    case = orc.make_case(orc.SCRIPTS["diff_3D_nonlin_perf"], 21, 21, 21)
    ref = orc.initial_condition(3, case, amp=2.0, bval=1.0)
    assert np.all(np.abs(H0 - ref) <= np.spacing(ref))
    assert np.all(H0[0] == 1.0) and np.all(H0[:, -1] == 1.0) and H0[10, 10, 10] == pytest.approx(3.0, rel=1e-2)


@pytest.mark.parametrize("script,kind", [("diff_1D_linstep", 0), ("diff_2D_linstep", 0),
                                         ("diff_3D_linstep_multixpu", 1)])
def test_step_field_matches_oracle(script, kind, pkg, orc):
    nx, ny, nz = _script_dims(script)
    D = pkg.step_field(pkg.SCRIPTS[script], nx, ny, nz)
    case = orc.make_case(orc.SCRIPTS[script], nx, ny, nz)
    assert np.array_equal(D, orc.step_field(kind, case))


def test_half_step_field_matches_oracle(pkg, orc):
    D = pkg.step_field(pkg.SCRIPTS["diff_3D_linstep_multixpu"], 17, 9, 9, kind="half_x")
    case = orc.make_case(orc.SCRIPTS["diff_3D_linstep_multixpu"], 17, 9, 9)
    assert np.array_equal(D, orc.step_field(2, case))


def test_field_shapes(pkg):
    from importlib import import_module
    abi = import_module("ptd_amd._abi")
    p3 = pkg.make_params(pkg.SCRIPTS["diff_3D_nonlin_perf"], 7, 6, 5)
    assert pkg.field_shape(p3, abi.F_H) == (7, 6, 5)
    assert pkg.field_shape(p3, abi.F_QX) == (6, 4, 3)
    assert pkg.field_shape(p3, abi.F_QY) == (5, 5, 3)
    assert pkg.field_shape(p3, abi.F_QZ) == (5, 4, 4)
    assert pkg.field_shape(p3, abi.F_RESH) == (5, 4, 3)
    p2 = pkg.make_params(pkg.SCRIPTS["diff_2D_lin"], 7, 6)
    assert pkg.field_shape(p2, abi.F_QX) == (6, 4) and pkg.field_shape(p2, abi.F_QY) == (5, 5)
    assert pkg.field_shape(p2, abi.F_QZ) is None
    p1 = pkg.make_params(pkg.SCRIPTS["diff_1D_lin"], 7)
    assert pkg.field_shape(p1, abi.F_QX) == (6,) and pkg.field_shape(p1, abi.F_RESH) == (5,)


def test_a_eff_matches_reference_definition(pkg):
    # (2*4+2)/1e9*nx*ny*nz*8 at 512^3 = 10.73741824 GB (output/out_diff_3D_nonlin_perf_ampere.txt:21)
    assert pkg.a_eff_gb(3, (512, 512, 512)) == pytest.approx(10.73741824, rel=1e-15)
    # (2*3+2)/1e9*nx*ny*8 at 8192^2 = 4.294967296 GB (output/out_diff_2D_nonlin_perf_ampere.txt:41)
    assert pkg.a_eff_gb(2, (8192, 8192)) == pytest.approx(4.294967296, rel=1e-15)
