"""Batches of independent solves (ptd_batch_*) against the same problems solved one at a time
(ptd_solve through a context): every member's H, Hold, fluxes, per-step iteration counts and errors
must be identical, bit for bit, for 2D and 3D members of every D type, in EXACT and FAST arithmetic,
over two consecutive solve calls (the batch replays one captured check-block graph many times; the
single-problem solves run eagerly). Batches run each member on the sweep its single solve picks: small
2D grids LDS-resident (also with PTD_LDS2D=0, which switches that off in both, so that the single-level
and two-level 2D batch sweeps are compared too), large 2D grids two-level, 3D grids three- and two-level."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (script, shape, amplitude, centre fraction)
MEMBERS = [
    ("diff_2D_lin", (40, 33), 1.0, (0.4, 0.6)),
    ("diff_2D_linstep", (64, 48), 1.5, (0.5, 0.5)),
    ("diff_2D_nonlin", (37, 37), 0.8, (0.6, 0.4)),
    ("diff_3D_lin_multixpu", (26, 20, 18), 1.2, (0.5, 0.4, 0.6)),
    ("diff_3D_linstep_multixpu", (30, 24, 20), 2.0, (0.5, 0.5, 0.5)),
    ("diff_3D_nonlin_multixpu", (28, 28, 28), 1.0, (0.45, 0.55, 0.5)),
    ("diff_3D_nonlin_multixpu", (67, 35, 21), 0.7, (0.3, 0.7, 0.5)),
    # LDS-resident in single solves (31 x 32 cores), two-level 2D sweep with PTD_LDS2D=0 (17 x 62 tiles)
    ("diff_2D_nonlin", (1000, 990), 1.0, (0.5, 0.45)),
    # more than 1024 cores (35 x 32): the two-level 2D sweep (19 x 63 tiles)
    ("diff_2D_lin", (1100, 1000), 0.9, (0.45, 0.5)),
    # the smallest grids: one inner cell / one inner row
    ("diff_3D_lin_multixpu", (3, 3, 3), 1.0, (0.5, 0.5, 0.5)),
    ("diff_2D_nonlin", (3, 4), 1.2, (0.5, 0.5)),
]


def _gauss(spec, shape, amp, centre):
    nd = len(shape)
    ls = [spec["lx"], spec.get("ly", 0.0), spec.get("lz", 0.0)][:nd]
    r2 = 0.0
    for d in range(nd):
        dx = ls[d] / shape[d]
        xc = dx / 2 + dx * np.arange(shape[d]) - centre[d] * ls[d]
        sh = [1] * nd
        sh[d] = shape[d]
        r2 = r2 + (xc ** 2).reshape(sh)
    return np.asfortranarray(amp * np.exp(-r2))


def _setup(pkg, arith):
    params, fields = [], []
    for script, shape, amp, centre in MEMBERS:
        spec = pkg.SCRIPTS[script]
        p = pkg.make_params(spec, *shape, arith=arith)
        H0 = _gauss(spec, shape, amp, centre)
        D = pkg.step_field(spec, *shape) if spec["dcase"] == pkg.STEP else None
        params.append(p)
        fields.append((H0, D))
    return params, fields


def _batch_vs_single(pkg, params, fields, nt=5):
    """Solve the members as one batch (two calls), then one at a time, and require equal bits."""
    with pkg.Batch(params) as b:
        for m, (H0, D) in enumerate(fields):
            b.set(m, "H", H0)
            b.set(m, "Hold", H0)
            if D is not None:
                b.set(m, "D", D)
        runs = []
        for call in range(2):
            it, er, lat, st, total = b.solve(nt)
            assert np.all(st == 0)
            assert np.all(lat > 0) and np.all(lat <= total * 1.01 + 1.0)
            runs.append((it, er, [b.get(m, "H") for m in range(len(params))],
                         [b.get(m, "Hold") for m in range(len(params))],
                         [[b.get(m, f) for f in ("qHx", "qHy", "qHz")[:params[m].ndim]] for m in range(len(params))]))
        tz, rows, ms = b.tiling()
        assert tz in (8, 16, 32) and rows in (8, 16)
    for m, (H0, D) in enumerate(fields):
        with pkg.Context(params[m]) as ctx:
            ctx.set("H", H0)
            ctx.set("Hold", H0)
            if D is not None:
                ctx.set("D", D)
            for call in range(2):
                it, er = ctx.solve(nt)
                bit, ber, bH, bHo, bq = runs[call]
                assert np.array_equal(bit[m], it), f"member {m} call {call}: {bit[m]} vs {it}"
                assert np.array_equal(ber[m], er), f"member {m} call {call}"
                assert np.array_equal(bH[m], ctx.get("H")), f"member {m} call {call}: H"
                assert np.array_equal(bHo[m], ctx.get("Hold")), f"member {m} call {call}: Hold"
                for a, f in enumerate(("qHx", "qHy", "qHz")[:params[m].ndim]):
                    assert np.array_equal(bq[m][a], ctx.get(f)), f"member {m} call {call}: {f}"


@pytest.mark.parametrize("lds", ["default", "off"])
@pytest.mark.parametrize("arith", [0, 1], ids=["exact", "fast"])
def test_batch_equals_single_solves(arith, lds, pkg, monkeypatch):
    if lds == "off":
        monkeypatch.setenv("PTD_LDS2D", "0")
    else:
        monkeypatch.delenv("PTD_LDS2D", raising=False)
    print(f"[gpu] batch vs single solves, arith={arith}, LDS-resident 2D {lds}", flush=True)
    params, fields = _setup(pkg, arith)
    _batch_vs_single(pkg, params, fields)


@pytest.mark.parametrize("nout", [6, 12, 14])
def test_batch_check_blocks_of_other_lengths(nout, pkg):
    """Check blocks of other lengths: the float64 FAST linear and step members take three-level passes,
    then two-level ones, and end every block on the checking two-level pass, in the batch and in their
    single solves alike (nout = 6 and 12 would otherwise end a single solve's block on a single-level
    checking sweep, whose residual partials are summed in another order)."""
    print(f"[gpu] batch vs single solves, nout={nout}", flush=True)
    params, fields = [], []
    for script, shape, amp, centre in MEMBERS[:6]:
        spec = pkg.SCRIPTS[script]
        p = pkg.make_params(spec, *shape, arith=1)
        p.nout = nout
        p.itmax = nout * 2000
        params.append(p)
        fields.append((_gauss(spec, shape, amp, centre),
                       pkg.step_field(spec, *shape) if spec["dcase"] == pkg.STEP else None))
    _batch_vs_single(pkg, params, fields)


def test_batch_workload_sample(pkg):
    """A reduced draw of the batched workload (small sizes): every member converges, counts are
    multiples of nout, latencies are ordered within the call."""
    from importlib import import_module
    B = import_module("ptd_amd.batch")
    w = B.workload(count=12, seed=3, sizes_2d=[128, 192], sizes_3d=[64, 96])
    params = [B.member_params(m) for m in w]
    print(f"[gpu] batch workload sample: {len(w)} members", flush=True)
    with pkg.Batch(params) as b:
        for i, m in enumerate(w):
            H0, D = B.member_fields(m)
            b.set(i, "H", H0)
            b.set(i, "Hold", H0)
            if D is not None:
                b.set(i, "D", D)
        it, er, lat, st, total = b.solve(5)
    assert np.all(st == 0)
    assert np.all(it > 0) and np.all(it % 10 == 0)
    assert np.all(er <= 1e-8)
    assert 0 < B.percentile(lat, 50) <= B.percentile(lat, 99) <= total * 1.01 + 1.0


def test_batch_solves_with_growing_step_counts(pkg):
    """A second solve call with more physical steps than the first reallocates the per-step outputs and
    re-captures the check-block graphs; every member still equals its single solves over the same calls."""
    print("[gpu] batch solve(2) then solve(4)", flush=True)
    members = [("diff_3D_nonlin_multixpu", (28, 24, 22), 1.0, (0.5, 0.5, 0.5)),
               ("diff_2D_linstep", (70, 60), 1.5, (0.4, 0.6))]
    params, fields = [], []
    for script, shape, amp, centre in members:
        spec = pkg.SCRIPTS[script]
        params.append(pkg.make_params(spec, *shape, arith=1))
        fields.append((_gauss(spec, shape, amp, centre),
                       pkg.step_field(spec, *shape) if spec["dcase"] == pkg.STEP else None))
    with pkg.Batch(params) as b:
        for m, (H0, D) in enumerate(fields):
            b.set(m, "H", H0)
            b.set(m, "Hold", H0)
            if D is not None:
                b.set(m, "D", D)
        runs = [b.solve(nt)[:2] for nt in (2, 4)]
        Hb = [b.get(m, "H") for m in range(len(params))]
    for m, (H0, D) in enumerate(fields):
        with pkg.Context(params[m]) as ctx:
            ctx.set("H", H0)
            ctx.set("Hold", H0)
            if D is not None:
                ctx.set("D", D)
            for call, nt in enumerate((2, 4)):
                it, er = ctx.solve(nt)
                assert np.array_equal(runs[call][0][m], it) and np.array_equal(runs[call][1][m], er)
            assert np.array_equal(Hb[m], ctx.get("H"))
