"""GPU parity of the variants the benchmark runs, against the CPU oracle.

The benchmark times FAST arithmetic (the rearranged update rules with Newton-refined hardware
reciprocals); these tests compare exactly those kernels with the oracle, which evaluates the reference's
expressions in their own order:
 - 2D FAST sweeps, single-level and two-level, at 64^2 and 1000x990: relative max-norm <= 1e-13 after 5
   sweeps from a random state (a few ulp per sweep);
 - 2D FAST converged solves (single-level and two-level): per-step iteration counts within one nout of
   the oracle's, final H within 1e-10 relative max-norm;
 - the benchmarked 3D sweep at 511^3 (FAST) and 255^3 (FAST against the oracle, <= 1e-13);
 - the step field of the 3D benchmark configuration (D = 1e-4 where the cell centre x < lx/2) at 511^3,
   and at 255^3 against the oracle (EXACT bit for bit, FAST <= 1e-13);
 - every member of a mixed batch against the oracle (EXACT bit for bit, FAST <= 1e-10 with counts within
   one nout), and the full 96-member batched workload, with a sample of its members against the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NOUT = 10


def _relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module", autouse=True)
def _device(pkg):
    lib = pkg.load()
    if lib.ptd_device_check(0, None, 0) != 0:
        pytest.fail(f"no gfx950 device: {lib.ptd_last_error()}")


def _random_state(pkg, orc, script, shape, seed):
    rng = np.random.default_rng(seed)
    spec = pkg.SCRIPTS[script]
    nx, ny, nz = (list(shape) + [1, 1])[:3]
    params = pkg.make_params(spec, nx, ny, nz, arith=1)
    case = orc.make_case(orc.SCRIPTS[script], nx, ny, nz)
    H = np.asfortranarray(rng.uniform(1.0, 1.2, size=shape))
    Hold = np.asfortranarray(rng.uniform(1.0, 1.2, size=shape))
    D = np.asfortranarray(np.where(rng.uniform(size=shape) < 0.5, 1e-4, 1.0)) if spec["dcase"] == pkg.STEP else None
    q = {}
    for name, fid in (("qHx", 3), ("qHy", 4), ("qHz", 5))[:len(shape)]:
        q[name] = np.asfortranarray(0.05 * rng.normal(size=pkg.field_shape(params, fid)))
    return params, case, H, Hold, D, q


def _oracle_iterate(orc, pkg, params, case, H, Hold, D, q, niter):
    dtau = None
    if D is not None:
        dtau = np.zeros(pkg.field_shape(params, 10), order="F")
        orc.lib().ptdo_iter_params(orc.ctypes.byref(case), orc._p(D), orc._p(dtau))
    Hr = H.copy(order="F")
    qr = {k: v.copy(order="F") for k, v in q.items()}
    orc.lib().ptdo_iterate(orc.ctypes.byref(case), orc._p(Hr), orc._p(Hold), orc._p(D), orc._p(dtau),
                           orc._p(qr.get("qHx")), orc._p(qr.get("qHy")), orc._p(qr.get("qHz")), niter)
    return Hr, qr


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


# --------------------------------------------------------------------------------------------------
# 2D FAST sweeps and solves
# --------------------------------------------------------------------------------------------------

T2D_SCRIPTS = ["diff_2D_lin", "diff_2D_linstep", "diff_2D_nonlin", "diff_2D_nonlin_perf"]
# (shape, PTD_T2): single-level sweeps at 64^2 (the library's choice below 1024 tiles of 60 x 16 cells),
# two-level sweeps forced at 64^2, and the library's own two-level choice at 1000 x 990 (17 x 62 tiles)
SWEEP_MODES = [((64, 64), "0"), ((64, 64), "2"), ((1000, 990), "1")]


@pytest.mark.parametrize("shape,t2", SWEEP_MODES, ids=["64-single", "64-twolevel", "1000x990-twolevel"])
@pytest.mark.parametrize("script", T2D_SCRIPTS)
def test_fast_2d_sweeps_close_to_oracle(script, shape, t2, pkg, orc, monkeypatch):
    """5 FAST sweeps (two-level: two pairs and one single sweep) from a random state: H within 1e-13 and
    the fluxes within 1e-12 relative max-norm of the oracle's 5 iterations."""
    print(f"[gpu] fast 2D sweeps {script} {shape} PTD_T2={t2}", flush=True)
    monkeypatch.setenv("PTD_T2", t2)
    monkeypatch.setenv("PTD_LDS2D", "0")  # the sweeps named by PTD_T2 (the LDS iterations have their own tests)
    params, case, H, Hold, D, q = _random_state(pkg, orc, script, shape, seed=sum(shape) + 3)
    Hg, qg = _gpu_iterate(pkg, params, H, Hold, D, q, 5)
    Hr, qr = _oracle_iterate(orc, pkg, params, case, H, Hold, D, q, 5)
    assert np.all(np.isfinite(Hr))
    assert _relmax(Hg, Hr) <= 1e-13, _relmax(Hg, Hr)
    for k in q:
        assert _relmax(qg[k], qr[k]) <= 1e-12, (k, _relmax(qg[k], qr[k]))


@pytest.mark.parametrize("t2", ["0", "2"], ids=["single", "twolevel"])
@pytest.mark.parametrize("script", T2D_SCRIPTS)
def test_fast_2d_solve_matches_oracle(script, t2, pkg, orc, monkeypatch):
    """Converging 5-step solve at 96 x 80 with FAST sweeps (single-level, or two-level pairs whose last
    pair of a check block carries the residual): per-step iteration counts within one nout of the
    oracle's and the final H within 1e-10 relative max-norm."""
    shape = (96, 80)
    print(f"[gpu] fast 2D solve {script} {shape} PTD_T2={t2}", flush=True)
    monkeypatch.setenv("PTD_T2", t2)
    monkeypatch.setenv("PTD_LDS2D", "0")
    over_o = dict(nout=NOUT, itmax=100000) if "perf" in script else {}
    over_g = dict(nout=NOUT, itMax=1e5) if "perf" in script else {}
    ref = orc.run_script(script, *shape, **over_o)
    res = pkg.run_script(script, *shape, H0=ref["H0"], D=ref["D"], arith=1, **over_g)
    assert all(abs(int(a) - int(b)) <= NOUT for a, b in zip(res.iters, ref["iters"])), \
        (list(res.iters), list(ref["iters"]))
    assert _relmax(res.H, ref["H"]) <= 1e-10, _relmax(res.H, ref["H"])


# --------------------------------------------------------------------------------------------------
# the benchmarked 3D sweep at full size, and the configuration's step field
# --------------------------------------------------------------------------------------------------

def _boundaries_unchanged(H, H0):
    for d in range(3):
        for edge in (0, H.shape[d] - 1):
            sl = [slice(None)] * 3
            sl[d] = edge
            assert np.array_equal(H[tuple(sl)], H0[tuple(sl)])


def test_fast_full_size_sweeps(pkg, orc):
    """The benchmarked sweep (FAST, two-level, 32-plane z chunks) at 511^3 on the benchmark's initial
    condition: finite, boundaries unchanged, x<->y symmetric to rounding (the FAST divergence sums the
    axes with fused multiply-adds, so the symmetry is not bitwise); and 4 FAST sweeps at 255^3 within
    1e-13 of the oracle."""
    spec = pkg.SCRIPTS["diff_3D_nonlin_perf"]
    n = 511
    print(f"[gpu] fast full size {n}^3", flush=True)
    params = pkg.make_params(spec, n, n, n, arith=1)
    H0 = pkg.initial_condition(spec, n, n, n, kind="config")
    with pkg.Context(params) as ctx:
        ctx.set("H", H0)
        ctx.set("Hold", H0)
        ctx.iterate(20)
        H = ctx.get("H")
    assert np.all(np.isfinite(H))
    _boundaries_unchanged(H, H0)
    assert _relmax(H, np.swapaxes(H, 0, 1)) <= 1e-14
    assert _relmax(H, H0) > 1e-6  # the sweeps moved the field
    del H, H0
    n = 255
    params = pkg.make_params(spec, n, n, n, arith=1)
    case = orc.make_case(orc.SCRIPTS["diff_3D_nonlin_perf"], n, n, n)
    H0 = pkg.initial_condition(spec, n, n, n, kind="config")
    q = {k: np.zeros(pkg.field_shape(params, f), order="F") for k, f in (("qHx", 3), ("qHy", 4), ("qHz", 5))}
    Hg, qg = _gpu_iterate(pkg, params, H0, H0, None, q, 4)
    Hr, qr = _oracle_iterate(orc, pkg, params, case, H0, H0, None, q, 4)
    assert _relmax(Hg, Hr) <= 1e-13, _relmax(Hg, Hr)
    for k in q:
        assert _relmax(qg[k], qr[k]) <= 1e-12, k


@pytest.mark.parametrize("arith", [0, 1], ids=["exact", "fast"])
def test_half_x_step_field(arith, pkg, orc):
    """The 3D benchmark configuration's step case, D = 1e-4 where the cell-centre x < lx/2 and 1 elsewhere,
    on the config IC: at 511^3 the sweeps stay finite, keep the boundaries and are y<->z symmetric to
    rounding (D depends on x only); at 255^3, 4 sweeps equal the oracle's bit for bit in EXACT and within
    1e-13 in FAST."""
    spec = pkg.SCRIPTS["diff_3D_linstep_multixpu"]
    n = 511
    print(f"[gpu] half-x step field {n}^3 arith={arith}", flush=True)
    params = pkg.make_params(spec, n, n, n, arith=arith)
    H0 = pkg.initial_condition(spec, n, n, n, kind="config")
    D = pkg.step_field(spec, n, n, n, kind="half_x")
    assert D[n // 2 - 1, 0, 0] == 1e-4 and D[n // 2 + 1, 0, 0] == 1.0
    with pkg.Context(params) as ctx:
        ctx.set("H", H0)
        ctx.set("Hold", H0)
        ctx.set("D", D)
        ctx.iterate(20)
        H = ctx.get("H")
    assert np.all(np.isfinite(H))
    _boundaries_unchanged(H, H0)
    assert _relmax(H, np.swapaxes(H, 1, 2)) <= 1e-14
    del H, H0, D
    n = 255
    params = pkg.make_params(spec, n, n, n, arith=arith)
    case = orc.make_case(orc.SCRIPTS["diff_3D_linstep_multixpu"], n, n, n)
    H0 = pkg.initial_condition(spec, n, n, n, kind="config")
    D = pkg.step_field(spec, n, n, n, kind="half_x")
    Do = orc.step_field(2, case)  # the oracle's own restatement of the same field
    assert np.array_equal(D, Do)
    q = {k: np.zeros(pkg.field_shape(params, f), order="F") for k, f in (("qHx", 3), ("qHy", 4), ("qHz", 5))}
    Hg, qg = _gpu_iterate(pkg, params, H0, H0, D, q, 4)
    Hr, qr = _oracle_iterate(orc, pkg, params, case, H0, H0, D, q, 4)
    if arith == 0:
        assert np.array_equal(Hg, Hr), _relmax(Hg, Hr)
        for k in q:
            assert np.array_equal(qg[k], qr[k]), k
    else:
        assert _relmax(Hg, Hr) <= 1e-13, _relmax(Hg, Hr)
        for k in q:
            assert _relmax(qg[k], qr[k]) <= 1e-12, k


# --------------------------------------------------------------------------------------------------
# batches against the oracle
# --------------------------------------------------------------------------------------------------

# (script, shape, amplitude, centre fraction): 2D and 3D members of every D type, the smallest grids
BATCH_MEMBERS = [
    ("diff_2D_lin", (40, 33), 1.0, (0.4, 0.6)),
    ("diff_2D_linstep", (64, 48), 1.5, (0.5, 0.5)),
    ("diff_2D_nonlin", (37, 37), 0.8, (0.6, 0.4)),
    ("diff_3D_lin_multixpu", (26, 20, 18), 1.2, (0.5, 0.4, 0.6)),
    ("diff_3D_linstep_multixpu", (30, 24, 20), 2.0, (0.5, 0.5, 0.5)),
    ("diff_3D_nonlin_multixpu", (28, 28, 28), 1.0, (0.45, 0.55, 0.5)),
    ("diff_3D_nonlin_multixpu", (67, 35, 21), 0.7, (0.3, 0.7, 0.5)),
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


def _oracle_solve(orc, script, shape, H0, D, nt=None):
    spec = dict(orc.SCRIPTS[script])
    case = orc.make_case(spec, *shape) if nt is None else orc.make_case(spec, *shape, nt=nt)
    H, Hold = H0.copy(order="F"), H0.copy(order="F")
    iters, errs, st = orc.solve(case, H, Hold, D)
    assert st == 0
    return H, iters, errs


def _check_member(arith, H, it, Hr, itr, label):
    if arith == 0:
        assert list(it) == list(itr), (label, list(it), list(itr))
        assert np.array_equal(H, Hr), (label, _relmax(H, Hr))
    else:
        assert all(abs(int(a) - int(b)) <= NOUT for a, b in zip(it, itr)), (label, list(it), list(itr))
        assert _relmax(H, Hr) <= 1e-10, (label, _relmax(H, Hr))


@pytest.mark.parametrize("arith", [0, 1], ids=["exact", "fast"])
def test_batch_members_match_oracle(arith, pkg, orc):
    """Every member of a mixed batch (2D / 3D, lin / step / H^3, down to one inner cell) against the oracle's
    solve of the same problem: EXACT bit for bit with equal per-step counts; FAST within 1e-10 with counts
    within one nout."""
    print(f"[gpu] batch members vs oracle, arith={arith}", flush=True)
    nt = 5
    params, fields = [], []
    for script, shape, amp, centre in BATCH_MEMBERS:
        spec = pkg.SCRIPTS[script]
        params.append(pkg.make_params(spec, *shape, arith=arith))
        fields.append((_gauss(spec, shape, amp, centre),
                       pkg.step_field(spec, *shape) if spec["dcase"] == pkg.STEP else None))
    with pkg.Batch(params) as b:
        for m, (H0, D) in enumerate(fields):
            b.set(m, "H", H0)
            b.set(m, "Hold", H0)
            if D is not None:
                b.set(m, "D", D)
        it, er, lat, st, total = b.solve(nt)
        Hb = [b.get(m, "H") for m in range(len(params))]
    assert np.all(st == 0)
    for m, (script, shape, _, _) in enumerate(BATCH_MEMBERS):
        H0, D = fields[m]
        Hr, itr, _ = _oracle_solve(orc, script, shape, H0, D)
        _check_member(arith, Hb[m], it[m], Hr, itr, f"member {m} {script} {shape}")


def test_batch_workload_96(pkg, orc):
    """The batched workload as the benchmark runs it: 96 members (PCG64 seed 0) of 2D {128..1024}^2 and 3D
    {64..256}^3 grids, lin / step / H^3, FAST, 5 physical steps in one ptd_batch_solve. Every member
    converges (status OK, error <= tol at every step, counts multiples of nout, latencies within the call),
    and the two smallest 2D and the two smallest 3D members agree with the oracle (counts within one nout,
    H within 1e-10)."""
    B = pkg.batch
    w = B.workload(count=96, seed=0)
    params = [B.member_params(m) for m in w]
    print(f"[gpu] batch workload: {len(w)} members", flush=True)
    with pkg.Batch(params) as b:
        for i, m in enumerate(w):
            H0, D = B.member_fields(m)
            b.set(i, "H", H0)
            b.set(i, "Hold", H0)
            if D is not None:
                b.set(i, "D", D)
        it, er, lat, st, total = b.solve(5)
        groups = b.group_times()
        order2 = sorted((m.n, i) for i, m in enumerate(w) if "2D" in m.script)[:2]
        order3 = sorted((m.n, i) for i, m in enumerate(w) if "3D" in m.script)[:2]
        sample = [i for _, i in order2 + order3]
        Hs = {i: b.get(i, "H") for i in sample}
    print(f"[gpu] batch workload: {total:.0f} ms, {int(it.sum())} PT iterations", flush=True)
    assert np.all(st == 0)
    assert np.all(er <= 1e-8)
    assert np.all(it > 0) and np.all(it % NOUT == 0)
    assert np.all(lat > 0) and np.all(lat <= total * 1.01 + 1.0)
    print(f"[gpu] batch groups: {groups}", flush=True)
    assert sum(g["members"] for g in groups) == 96
    assert all(0 < g["ms"] <= total * 1.01 for g in groups)
    for i in sample:
        m = w[i]
        H0, D = B.member_fields(m)
        shape = H0.shape
        Hr, itr, _ = _oracle_solve(orc, m.script, shape, H0, D)
        print(f"[gpu]   member {i} {m.script} n={m.n}: {list(it[i])} vs oracle {list(itr)}", flush=True)
        _check_member(1, Hs[i], it[i], Hr, itr, f"member {i} {m.script} n={m.n}")
