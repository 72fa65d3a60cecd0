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


@pytest.mark.parametrize("name,expect_file,inner", [
    ("diff_1D_lin", "diff_1D_lin.mat", False),
    ("diff_2D_nonlin_perf", "diff_2D_nonlin_perf.mat", False),
    ("diff_3D_nonlin_multixpu", "diff_3D_nonlin.mat", True),
    ("diff_3D_linstep_multixpu", "diff_3D_linstep.mat", True),
    ("diff_3D_nonlin_perf", "diff_3D_nonlin_perf.mat", False),
    ("diff_3D_nonlin_multixpu_perf", "diff_3D_nonlin_multixpu_perf.mat", True),
])
def test_save_viz_mat_v5(name, expect_file, inner, tmp_path):
    """DO_SAVE_VIZ dumps (scripts/diff_1D_lin.jl:92, scripts/diff_3D/diff_3D_nonlin_multixpu.jl:137,
    scripts/diff_3D/diff_3D_nonlin_perf.jl:133): file name, variable names, inn(H) for the multi-xPU
    scripts, coordinate vectors as n x 1 columns; read back with scipy.io.loadmat."""
    import scipy.io
    from importlib import import_module
    scripts = import_module("ptd_amd.scripts")
    nd = scripts.SCRIPTS[name]["ndim"]
    n = [7, 6, 5][:nd]
    rng = np.random.default_rng(0)
    H = np.asfortranarray(rng.random(n))
    coords = tuple(scripts.linrange(0.5 / m, 10 - 0.5 / m, m) for m in n)
    r = scripts.ScriptResult(name=name, coords=coords, H=H, iters=np.zeros(1, np.int64), errs=np.zeros(1))
    path = scripts.save_viz(name, r, out_dir=str(tmp_path))
    assert path.endswith(expect_file)
    m = scipy.io.loadmat(path)
    want = H[(slice(1, -1),) * nd] if inner else H
    got = m[f"H_{nd}D"]
    assert np.array_equal(got.reshape(want.shape) if nd == 1 else got, want)
    for d, c in enumerate(("xc", "yc", "zc")[:nd]):
        v = m[f"{c}_{nd}D"]
        assert v.shape == (n[d], 1)
        assert np.array_equal(v[:, 0], coords[d])


def test_save_viz_skips_reaction_scripts(tmp_path):
    from importlib import import_module
    scripts = import_module("ptd_amd.scripts")
    r = scripts.ScriptResult(name="diff_1D_react", coords=(np.zeros(4),), H=np.zeros(4),
                             iters=np.zeros(1, np.int64), errs=np.zeros(1))
    assert scripts.save_viz("diff_1D_react", r, out_dir=str(tmp_path)) is None


def test_batch_workload_draw(pkg):
    """The batched workload: 96 members from PCG64(0), reproducible, with every field inside the ranges
    the workload states (2D {128, ..., 1024}^2, 3D {64, ..., 256}^3, D from lin / step / H^3, amplitude
    U(0.5, 2), centre U(0.3, 0.7) per axis)."""
    B = pkg.batch
    w = B.workload(96, 0)
    assert [(m.script, m.n, m.amp, m.centre) for m in w] == [(m.script, m.n, m.amp, m.centre) for m in B.workload(96, 0)]
    assert len(w) == 96
    kinds = {(pkg.SCRIPTS[m.script]["ndim"], pkg.SCRIPTS[m.script]["dcase"]) for m in w}
    assert kinds == {(d, c) for d in (2, 3) for c in (pkg.LIN, pkg.STEP, pkg.NONLIN)}
    for m in w:
        nd = pkg.SCRIPTS[m.script]["ndim"]
        assert m.n in (B.SIZES_2D if nd == 2 else B.SIZES_3D)
        assert 0.5 <= m.amp < 2.0 and len(m.centre) == nd and all(0.3 <= c < 0.7 for c in m.centre)
    assert B.SIZES_2D == list(range(128, 1025, 64)) and B.SIZES_3D == list(range(64, 257, 32))


# Rows of the reference's own output files (data): the converging scripts' `nx [ny [nz]] ittot nt`
# (output/out_diff_1D_lin.txt:1, out_diff_2D_lin.txt:1, out_diff_3D_lin.txt:1) and the perf scripts'
# rows with T_eff (output/out_diff_2D_nonlin_perf_ampere.txt:1, out_diff_3D_nonlin_perf_ampere.txt:1,
# out_diff_3D_nonlin_multixpu_perf_ampere.txt:1).
REFERENCE_ROWS = {
    1: "32 200 5",
    2: "32 32 220 5",
    3: "62 62 62 430 5",
    "2D_perf": "32 32 200 0.003325939178466797 6.5536e-5 3.325939178466797e-5 1.9704509458351256",
    "3D_perf": "1 32 32 32 2000 0.038771867752075195 0.00262144 3.87718677520752e-5 67.61190914924886",
    "3D_multixpu_perf": "1 512 512 512 400 2.2332510948181152 10.73741824 0.011816143358826007 908.7075125895237",
}


def _shape(tok):
    return ["i" if re.fullmatch(r"-?\d+", t) else "f" for t in tok.split()]


def test_julia_float_reproduces_reference_rows():
    """julia_float writes every float of the reference's rows exactly as the file holds it."""
    from importlib import import_module
    S = import_module("ptd_amd.scripts")
    for row in REFERENCE_ROWS.values():
        for tok in row.split():
            if "." in tok or "e" in tok:
                assert S.julia_float(float(tok)) == tok, tok


@pytest.mark.parametrize("name", ["diff_1D_lin", "diff_1D_linstep", "diff_1D_nonlin", "diff_2D_lin",
                                  "diff_2D_linstep", "diff_2D_nonlin", "diff_2D_nonlin_perf", "diff_3D_lin_multixpu",
                                  "diff_3D_linstep_multixpu", "diff_3D_nonlin_multixpu", "diff_3D_nonlin_perf",
                                  "diff_3D_nonlin_multixpu_perf", "diff_1D_react", "diff_2D_react"])
def test_main_output_rows(name, tmp_path, monkeypatch, capsys):
    """main() with DO_SAVE=1 on a stubbed GPU run: the row lands in the file the reference appends to
    (the multi-xPU 3D converging scripts without their suffix), with the reference row's fields in order;
    the summary lines are the scripts' printf lines (the 1D linear one with its analytic error); the
    reaction scripts write no row."""
    from importlib import import_module
    S = import_module("ptd_amd.scripts")
    spec = S.SCRIPTS[name]
    nd = spec["ndim"]
    n = {1: [32], 2: [32, 32], 3: [32, 32, 32]}[nd]

    def fake_run(nm, nx, ny=None, nz=None, **kw):
        coords = tuple(S.linrange(10.0 / v / 2, 10.0 - 10.0 / v / 2, v) for v in n)
        H = np.asfortranarray(np.full(tuple(n), 0.5))
        iters = np.array([40, 30, 30, 30, 30][: 1 if spec.get("react") else (2 if spec.get("perf") else 5)])
        r = S.ScriptResult(nm, coords, H, iters, np.full(len(iters), 9e-9))
        if spec.get("perf"):
            r.t_toc, r.niter = 0.0123, int(iters[-1])
            r.A_eff = S.a_eff_gb(nd, n)
            r.t_it = r.t_toc / r.niter
            r.T_eff = r.A_eff / r.t_it
        return r

    monkeypatch.setattr(S, "run_script", fake_run)
    monkeypatch.chdir(tmp_path)
    for k, v in zip(("NX", "NY", "NZ"), n):
        monkeypatch.setenv(k, str(v))
    monkeypatch.setenv("DO_SAVE", "true")
    assert S.main([name]) == 0
    printed = capsys.readouterr().out.splitlines()
    files = sorted(p.name for p in (tmp_path / "output").glob("*")) if (tmp_path / "output").exists() else []
    if spec.get("react"):
        assert files == []
        assert printed[-1].startswith("nx = 32, ")
        return
    expect = {"diff_3D_lin_multixpu": "out_diff_3D_lin.txt", "diff_3D_linstep_multixpu": "out_diff_3D_linstep.txt",
              "diff_3D_nonlin_multixpu": "out_diff_3D_nonlin.txt"}.get(name, f"out_{name}.txt")
    assert files == [expect]
    row = (tmp_path / "output" / expect).read_text().splitlines()
    assert len(row) == 1
    ref = REFERENCE_ROWS["2D_perf" if name == "diff_2D_nonlin_perf" else "3D_perf" if name == "diff_3D_nonlin_perf"
                         else "3D_multixpu_perf" if name == "diff_3D_nonlin_multixpu_perf" else nd]
    assert _shape(row[0]) == _shape(ref), (row[0], ref)
    toks = row[0].split()
    if spec.get("perf"):
        assert toks[: len(toks) - 5][-nd:] == [str(v) for v in n]
    else:
        assert toks == [str(v) for v in n] + ["160", "5"]
    total = [ln for ln in printed if ln.startswith("Total time = ")]
    assert len(total) == 1 and total[0].endswith(" ")
    if spec.get("perf"):
        assert printed[0].startswith("PERF: Time = 0.012 sec, T_eff = ")
    if name == "diff_1D_lin":
        m = re.search(r"error vs analytic = ([0-9.e+-]+) $", total[0])
        assert m, total[0]
        xc = S.linrange(10.0 / 64, 10.0 - 10.0 / 64, 32)
        hana = 1 / np.sqrt(4 * 1.25) * np.exp(-(xc - 5.0) ** 2 / (4 * 1.25))
        assert m.group(1) == "%1.2e" % (np.linalg.norm(0.5 - hana) / np.sqrt(32))


def test_script_command_line_arguments(pkg, monkeypatch, capsys):
    """The scripts' command line (`python __graft_entry__.py script <name>`) refuses an unknown script and a
    PTD_ARITH other than exact / fast before it touches the device."""
    from importlib import import_module
    S = import_module("ptd_amd.scripts")
    assert S.main([]) == 2
    assert S.main(["diff_9D_lin"]) == 2
    assert "usage" in capsys.readouterr().err
    monkeypatch.setenv("PTD_ARITH", "fp16")
    assert S.main(["diff_2D_lin"]) == 2
    assert "PTD_ARITH" in capsys.readouterr().err
