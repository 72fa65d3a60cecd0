"""Host side of the reference scripts, running their PT loops on the GPU through libptd.

Each entry of SCRIPTS restates one script's "Physics"/"Numerics" block; `run_script(name, ...)` is the
equivalent of `include("scripts/<name>.jl")` with USE_RETURN=true: it builds the same initial
condition and D field on the host, derives the same scalars (Vpdtau, Re, theta_r_dtau, dtau_rho) in the
same operation order, runs the physical-time loop on the device and returns the coordinates and the
final H. `main()` reads the scripts' environment interface (NX, NY, NZ, DO_SAVE, ...) and prints and
appends the same lines the scripts do.
"""
import math
import os
import sys
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _abi
from .core import Context

LIN, STEP, NONLIN = _abi.LIN, _abi.STEP, _abi.NONLIN
PLAIN, PERF = _abi.PLAIN, _abi.PERF

# ic: "linrange_shift" (1D/2D: exp.(-(xc .- lx/2).^2 ...), the shifted coordinate being a LinRange),
#     "xc" (3D perf comprehension over xc[ix] - 0.5*lx), "xg" (3D multixpu: ImplicitGlobalGrid x_g)
# dstep: "index" (1D/2D D[1:ceil(n/2.2)] = D2), "xg" (3D: x_g < lx/2.2 or y_g < ly/2.2)
SCRIPTS = {
    "diff_1D_lin": dict(ndim=1, dcase=LIN, form=PLAIN, lx=10.0, D=1.0, ttot=1.0, dt=0.2, tol=1e-8,
                        itMax=1e5, nout=10, CFL=0.99, Resc=1.0, ic="linrange_shift",
                        src="scripts/diff_1D_lin.jl:34-50"),
    "diff_1D_linstep": dict(ndim=1, dcase=STEP, form=PLAIN, lx=10.0, D1=1.0, D2=1e-4, ttot=1.0, dt=0.2,
                            tol=1e-8, itMax=1e5, nout=10, CFL=0.99, Resc=1.0, ic="linrange_shift",
                            dstep="index", src="scripts/diff_1D_linstep.jl:45-61"),
    "diff_1D_nonlin": dict(ndim=1, dcase=NONLIN, form=PLAIN, lx=10.0, ttot=1.0, dt=0.2, tol=1e-8,
                           itMax=1e5, nout=10, CFL=1.0, Resc=1.0, ic="linrange_shift",
                           src="scripts/diff_1D_nonlin.jl:40-53"),
    "diff_1D_react": dict(ndim=1, dcase=LIN, form=PLAIN, lx=20.0, D=1.0, theta_kin=0.1, tol=1e-8,
                          itMax=1e5, nout=10, CFL=0.99, Resc=1.0, ic="linrange_shift", react=True,
                          src="scripts/diff_1D_react.jl:38-53"),
    "diff_2D_lin": dict(ndim=2, dcase=LIN, form=PLAIN, lx=10.0, ly=10.0, D=1.0, ttot=1.0, dt=0.2, tol=1e-8,
                        itMax=1e5, nout=10, CFL=1 / math.sqrt(2), Resc=1.0, ic="linrange_shift",
                        src="scripts/diff_2D_lin.jl:38-53"),
    "diff_2D_linstep": dict(ndim=2, dcase=STEP, form=PLAIN, lx=10.0, ly=10.0, D1=1.0, D2=1e-4, ttot=1.0,
                            dt=0.2, tol=1e-8, itMax=1e5, nout=10, CFL=1 / math.sqrt(2), Resc=1.0,
                            ic="linrange_shift", dstep="index", src="scripts/diff_2D_linstep.jl:55-83"),
    "diff_2D_nonlin": dict(ndim=2, dcase=NONLIN, form=PLAIN, lx=10.0, ly=10.0, ttot=1.0, dt=0.2, tol=1e-8,
                           itMax=1e5, nout=10, CFL=1 / math.sqrt(2), Resc=1 / 1.2, ic="linrange_shift",
                           src="scripts/diff_2D_nonlin.jl:46-58"),
    "diff_2D_nonlin_perf": dict(ndim=2, dcase=NONLIN, form=PERF, lx=10.0, ly=10.0, ttot=0.4, dt=0.2,
                                tol=1e-8, itMax=1e2, nout=2000, CFL=1 / math.sqrt(2), Resc=1 / 1.2,
                                ic="linrange_shift", perf=True, src="scripts/diff_2D_nonlin_perf.jl:56-71"),
    "diff_2D_react": dict(ndim=2, dcase=LIN, form=PLAIN, lx=10.0, ly=10.0, D=1.0, theta_kin=0.1, tol=1e-8,
                          itMax=1e5, nout=10, CFL=1 / math.sqrt(2), Resc=1.0, ic="linrange_shift", react=True,
                          src="scripts/diff_2D_react.jl:38-53"),
    "diff_3D_lin_multixpu": dict(ndim=3, dcase=LIN, form=PLAIN, lx=10.0, ly=10.0, lz=10.0, D=1.0, ttot=1.0,
                                 dt=0.2, tol=1e-8, itMax=1e5, nout=10, CFL=1 / math.sqrt(3), Resc=1.0,
                                 ic="xg", src="scripts/diff_3D/diff_3D_lin_multixpu.jl:45-62"),
    "diff_3D_linstep_multixpu": dict(ndim=3, dcase=STEP, form=PLAIN, lx=10.0, ly=10.0, lz=10.0, D1=1.0,
                                     D2=1e-4, ttot=1.0, dt=0.2, tol=1e-8, itMax=1e5, nout=10,
                                     CFL=1 / math.sqrt(3), Resc=1.0, ic="xg", dstep="xg",
                                     src="scripts/diff_3D/diff_3D_linstep_multixpu.jl:67-104"),
    "diff_3D_nonlin_multixpu": dict(ndim=3, dcase=NONLIN, form=PLAIN, lx=10.0, ly=10.0, lz=10.0, ttot=1.0,
                                    dt=0.2, tol=1e-8, itMax=1e5, nout=10, CFL=1 / math.sqrt(3),
                                    Resc=1 / 1.2, ic="xg", src="scripts/diff_3D/diff_3D_nonlin_multixpu.jl:57-75"),
    "diff_3D_nonlin_perf": dict(ndim=3, dcase=NONLIN, form=PERF, lx=10.0, ly=10.0, lz=10.0, ttot=0.4,
                                dt=0.2, tol=1e-8, itMax=1e3, nout=2000, CFL=1 / math.sqrt(3), Resc=1 / 1.2,
                                ic="xc", perf=True, src="scripts/diff_3D/diff_3D_nonlin_perf.jl:61-78"),
    "diff_3D_nonlin_multixpu_perf": dict(ndim=3, dcase=NONLIN, form=PERF, lx=10.0, ly=10.0, lz=10.0,
                                         ttot=0.4, dt=0.2, tol=1e-8, itMax=1e3, nout=2000,
                                         CFL=1 / math.sqrt(3), Resc=1 / 1.2, ic="xg", perf=True,
                                         src="scripts/diff_3D/diff_3D_nonlin_multixpu_perf.jl:61-80"),
}


def linrange(a, b, n):
    """Julia LinRange(a, b, n) elements: (1-t)*a + t*b with t = (i-1)/(n-1) (Base lerpi)."""
    t = np.arange(n, dtype=np.float64) / float(max(n - 1, 1))
    return (1.0 - t) * a + t * b


def make_params(spec, nx, ny=1, nz=1, precision=_abi.FP64, form=None, stop_norm=_abi.NORM_RMS,
                norm_len=0.0, arith=_abi.ARITH_EXACT):
    """ptd_params from a script's constants, derived as in its "Derived numerics" block
    (e.g. scripts/diff_3D/diff_3D_nonlin_perf.jl:73-78, scripts/diff_2D_lin.jl:48-53)."""
    nd = spec["ndim"]
    lx, ly, lz = spec["lx"], spec.get("ly", 1.0), spec.get("lz", 1.0)
    dt = spec["theta_kin"] if spec.get("react") else spec["dt"]
    dx = lx / nx
    dy = ly / ny if nd >= 2 else 1.0
    dz = lz / nz if nd >= 3 else 1.0
    if nd == 1:
        vpdt, max_l = spec["CFL"] * dx, lx
    elif nd == 2:
        vpdt, max_l = spec["CFL"] * min(dx, dy), max(lx, ly)
    else:
        vpdt, max_l = spec["CFL"] * min(dx, dy, dz), max(lx, ly, lz)
    max_l2 = max_l * max_l
    d0 = spec.get("D", 1.0)
    d_re = max(spec.get("D1", 1.0), spec.get("D2", 1.0)) if spec["dcase"] == STEP else d0
    re = math.pi + math.sqrt(math.pi * math.pi + (max_l2 / d_re) / dt)   # Re (lin/step), Da (react)
    p = _abi.Params()
    p.ndim, p.dcase = nd, spec["dcase"]
    p.form = spec["form"] if form is None else form
    p.precision = precision
    p.nx, p.ny, p.nz = nx, (ny if nd >= 2 else 1), (nz if nd >= 3 else 1)
    p.itmax = int(spec["itMax"])
    p.nout = int(spec["nout"])
    p.stop_norm = stop_norm
    p.arith = arith
    p.dx, p.dy, p.dz, p.dt = dx, dy, dz, dt
    p.vpdt, p.max_l, p.max_l2 = vpdt, max_l, max_l2
    p.resc, p.d0, p.re = spec["Resc"], d0, re
    p.theta_r = max_l / vpdt / re                  # theta_r_dtau = max(lx,...)/Vpdtau/Re
    p.dtau_rho = vpdt * max_l / d0 / re            # dtau_rho = Vpdtau*max(lx,...)/D/Re
    p.tol = spec["tol"]
    p.norm_len = norm_len
    return p


def _axis(kind, n, l):
    dxl = l / n
    if kind == "linrange_shift":
        # (LinRange(dx/2, l-dx/2, n) .- l/2) is itself a LinRange (range broadcasting)
        return linrange(dxl / 2 - l / 2, (l - dxl / 2) - l / 2, n)
    if kind == "xc":
        return linrange(dxl / 2, l - dxl / 2, n) - 0.5 * l
    if kind == "xg":
        # x_g(ix, dx, A) - 0.5*l + dx/2 with x_g = (ix-1)*dx on a single rank
        return (np.arange(n, dtype=np.float64) * dxl - 0.5 * l) + dxl / 2
    raise ValueError(kind)


def initial_condition(spec, nx, ny=1, nz=1, kind=None, amp=2.0, bval=1.0):
    """H0 as the script builds it (column-major (nx[,ny[,nz]])).

    kind "config": 1 + amp*exp(-r^2) on the inner cells, bval on the boundary cells (benchmark setup).
    """
    nd = spec["ndim"]
    kind = kind or spec["ic"]
    ak = "xc" if kind == "config" else kind
    n = [nx, ny, nz][:nd]
    ls = [spec["lx"], spec.get("ly", 0.0), spec.get("lz", 0.0)][:nd]
    sq = [(lambda a: a * a)(_axis(ak, n[d], ls[d])) for d in range(nd)]
    # build in C order (z, y, x) and transpose to a column-major (x, y, z) view
    e = -sq[0]
    if nd >= 2:
        e = e[None, :] - sq[1][:, None]
    if nd >= 3:
        e = e[None, :, :] - sq[2][:, None, None]
    H0 = np.exp(e).T
    if kind == "config":
        H0 = 1.0 + amp * H0
        idx = [slice(None)] * nd
        for d in range(nd):
            for edge in (0, n[d] - 1):
                idx2 = list(idx)
                idx2[d] = edge
                H0[tuple(idx2)] = bval
    return np.asfortranarray(H0)


def step_field(spec, nx, ny=1, nz=1, kind=None, d1=None, d2=None):
    """D of the step scripts: D2 on the low side, D1 elsewhere."""
    nd = spec["ndim"]
    kind = kind or spec.get("dstep", "index")
    d1 = spec.get("D1", 1.0) if d1 is None else d1
    d2 = spec.get("D2", 1e-4) if d2 is None else d2
    n = [nx, ny, nz][:nd]
    ls = [spec["lx"], spec.get("ly", 0.0), spec.get("lz", 0.0)][:nd]
    low = np.zeros(tuple(n), dtype=bool, order="F")
    if kind == "index":          # D[1:Int(ceil(nx/2.2)),:] .= D2; D[:,1:Int(ceil(ny/2.2))] .= D2
        for d in range(min(nd, 2)):
            c = int(math.ceil(n[d] / 2.2))
            sl = [slice(None)] * nd
            sl[d] = slice(0, c)
            low[tuple(sl)] = True
    elif kind == "xg":           # D[x_g .< lx/2.2] .= D2; D[y_g .< ly/2.2] .= D2
        for d in range(min(nd, 2)):
            xg = np.arange(n[d], dtype=np.float64) * (ls[d] / n[d])
            sh = [1] * nd
            sh[d] = n[d]
            low |= (xg < ls[d] / 2.2).reshape(sh, order="F")
    elif kind == "half_x":       # D2 where the cell-centre x < lx/2
        dxl = ls[0] / n[0]
        xc = linrange(dxl / 2, ls[0] - dxl / 2, n[0])
        sh = [1] * nd
        sh[0] = n[0]
        low |= (xc < ls[0] / 2).reshape(sh, order="F")
    else:
        raise ValueError(kind)
    return np.asfortranarray(np.where(low, d2, d1))


@dataclass
class ScriptResult:
    name: str
    coords: tuple
    H: np.ndarray
    iters: np.ndarray
    errs: np.ndarray
    t_toc: Optional[float] = None     # wall time of the timed physical step (perf scripts)
    niter: Optional[int] = None       # iterations in the timed step
    A_eff: Optional[float] = None     # GB per iteration, reference definition
    t_it: Optional[float] = None
    T_eff: Optional[float] = None
    extra: dict = field(default_factory=dict)

    @property
    def ittot(self):
        return int(self.iters.sum())


def a_eff_gb(ndim, n, esz=8):
    """Effective memory access per iteration [GB]: (2*(ndim+1)+2)*prod(n)*sizeof
    (scripts/diff_3D/diff_3D_nonlin_perf.jl:114, scripts/diff_2D_nonlin_perf.jl:107)."""
    cells = 1
    for v in n:
        cells *= v
    return (2 * (ndim + 1) + 2) / 1e9 * cells * esz


def run_script(name, nx, ny=None, nz=None, precision=_abi.FP64, device=0, ic=None, amp=2.0, bval=1.0,
               dstep=None, H0=None, D=None, **over):
    """include("scripts/<name>.jl") with USE_RETURN=true, on the GPU. Returns a ScriptResult.

    H0 / D, when given, replace the script's initial condition / step field (same shape)."""
    spec = dict(SCRIPTS[name])
    spec.update(over)
    nd = spec["ndim"]
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    n = [nx, ny, nz][:nd]
    params = make_params(spec, nx, ny, nz, precision=precision, norm_len=spec.get("norm_len", 0.0),
                         stop_norm=spec.get("stop_norm", _abi.NORM_RMS),
                         arith=spec.get("arith", _abi.ARITH_EXACT))
    dxs = [spec[k] / n[d] for d, k in enumerate(("lx", "ly", "lz")[:nd])]
    coords = tuple(linrange(dxs[d] / 2, spec[("lx", "ly", "lz")[d]] - dxs[d] / 2, n[d]) for d in range(nd))
    if H0 is None:
        H0 = initial_condition(spec, nx, ny, nz, kind=ic, amp=amp, bval=bval)
    nt = 1 if spec.get("react") else int(math.ceil(spec["ttot"] / spec["dt"]))
    with Context(params, device) as ctx:
        if spec.get("react"):
            ctx.set("H", np.zeros_like(H0))        # H = @zeros(nx): boundaries held at 0
            ctx.set("Hold", H0)                    # Heq
        else:
            ctx.set("H", H0)
            ctx.set("Hold", H0)
        if spec["dcase"] == STEP:
            ctx.set("D", step_field(spec, nx, ny, nz, kind=dstep) if D is None else D)
            ctx.compute_iter_params()
        iters = np.zeros(nt, dtype=np.int64)
        errs = np.zeros(nt)
        t_toc = None
        niter = None
        for it in range(nt):
            t0 = time.perf_counter()
            iters[it], rms, mx = ctx.pt_solve(raise_nan=False)
            t1 = time.perf_counter()
            errs[it] = mx if params.stop_norm == _abi.NORM_MAX else rms
            if it == 1:                             # the reference times physical step 2 only
                t_toc, niter = t1 - t0, int(iters[it])
            if not spec.get("react"):
                ctx.assign()                        # Hold .= H
            if math.isnan(errs[it]):
                raise _abi.PTDError(_abi.ERR_NAN, "NaN")
        H = ctx.get("H")
    res = ScriptResult(name, coords, H, iters, errs, t_toc=t_toc, niter=niter)
    if spec.get("perf") and niter:
        esz = 4 if precision == _abi.FP32 else 8
        res.A_eff = a_eff_gb(nd, n, esz)
        res.t_it = t_toc / niter
        res.T_eff = res.A_eff / res.t_it
    return res


def _env_bool(key, default):
    v = os.environ.get(key)
    return default if v is None else v.strip().lower() in ("1", "true", "yes")


def viz_file_name(name):
    """File name of the script's MAT dump: the multi-xPU 3D scripts drop the `_multixpu` suffix
    (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:137), the others keep their name
    (scripts/diff_1D_lin.jl:92, scripts/diff_3D/diff_3D_nonlin_perf.jl:133)."""
    if name.endswith("_multixpu"):
        name = name[: -len("_multixpu")]
    return name + ".mat"


def save_viz(name, result, out_dir="out_visu"):
    """The scripts' DO_SAVE_VIZ dump: a compressed MAT v5 file with H_<n>D and the cell-centre
    coordinates xc_<n>D, yc_<n>D, zc_<n>D (vectors as n x 1 columns, as MAT.jl writes Julia vectors).
    The multi-xPU scripts save the gathered inner array inn(H) (H_v, one rank:
    scripts/diff_3D/diff_3D_nonlin_multixpu.jl:92-94,123), the others the whole H. The reaction scripts
    save nothing."""
    import scipy.io

    spec = SCRIPTS[name]
    if spec.get("react"):
        return None
    nd = spec["ndim"]
    H = result.H
    if name.endswith("_multixpu") or name.endswith("_multixpu_perf"):
        H = H[(slice(1, -1),) * nd]
    data = {f"H_{nd}D": np.asfortranarray(H)}
    for d, c in enumerate(("xc", "yc", "zc")[:nd]):
        data[f"{c}_{nd}D"] = np.asarray(result.coords[d], dtype=np.float64)
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, viz_file_name(name))
    scipy.io.savemat(path, data, do_compression=True, format="5", oned_as="column")
    return path


def text_file_name(name):
    """File the script's DO_SAVE row is appended to: the multi-xPU 3D converging scripts drop the
    `_multixpu` suffix (scripts/diff_3D/diff_3D_lin_multixpu.jl:119, diff_3D_linstep_multixpu.jl:151,
    diff_3D_nonlin_multixpu.jl:131), the perf scripts keep their full name
    (scripts/diff_3D/diff_3D_nonlin_multixpu_perf.jl:145, diff_3D_nonlin_perf.jl:127), the 1D / 2D scripts
    use theirs (scripts/diff_1D_lin.jl:86, scripts/diff_2D_lin.jl:88)."""
    if name.endswith("_multixpu"):
        name = name[: -len("_multixpu")]
    return f"out_{name}.txt"


def julia_float(x):
    """A Float64 as Julia's string interpolation writes it: the shortest round-trip digits, `1.0e-5`-style
    exponents below 1e-4."""
    r = repr(float(x))
    if "e" not in r:
        return r
    mant, exp = r.split("e")
    if "." not in mant:
        mant += ".0"
    sign = "-" if exp.startswith("-") else ""
    return f"{mant}e{sign}{exp.lstrip('+-').lstrip('0') or '0'}"


def output_row(name, result, n):
    """The line the script appends with DO_SAVE:
    converging 1D / 2D / 3D: `nx [ny [nz]] ittot nt`  (scripts/diff_1D_lin.jl:87, diff_2D_lin.jl:89,
                                                       diff_3D/diff_3D_nonlin_multixpu.jl:132)
    2D perf:    `nx ny ittot t_toc A_eff t_it T_eff`    (scripts/diff_2D_nonlin_perf.jl:117)
    3D perf:    `1 nx ny nz ittot t_toc A_eff t_it T_eff` (scripts/diff_3D/diff_3D_nonlin_perf.jl:128)
    3D multi-xPU perf: `nprocs nx_g ny_g nz_g ittot ...` with nprocs = 1 here
                                                      (scripts/diff_3D/diff_3D_nonlin_multixpu_perf.jl:146)."""
    spec = SCRIPTS[name]
    dims = " ".join(str(int(v)) for v in n)
    if spec.get("perf"):
        prefix = "1 " if spec["ndim"] == 3 else ""
        vals = " ".join(julia_float(v) for v in (result.t_toc, result.A_eff, result.t_it, result.T_eff))
        return f"{prefix}{dims} {result.ittot} {vals}"
    return f"{dims} {result.ittot} {len(result.iters)}"


def analytic_error_1d(result, spec):
    """norm(H - Hana)/sqrt(nx) with Hana = 1/sqrt(4(ttot + 1/4)) exp(-(xc - lx/2)^2/(4 D (ttot + 1/4))), the
    linear 1D script's error against the analytic solution (scripts/diff_1D_lin.jl:80-81)."""
    xc = result.coords[0]
    ttot, lx, D = spec["ttot"], spec["lx"], spec["D"]
    hana = 1 / math.sqrt(4 * (ttot + 1 / 4)) * np.exp(-(xc - lx / 2) ** 2 / (4 * D * (ttot + 1 / 4)))
    return float(np.linalg.norm(result.H - hana) / math.sqrt(len(xc)))


def _round_sig(x, sig):
    return float(f"{x:.{sig}g}")


def summary_lines(name, result, n):
    """The lines the script prints after its loop (scripts/diff_3D/diff_3D_nonlin_perf.jl:117-118,
    scripts/diff_1D_lin.jl:81, scripts/diff_1D_react.jl:70, scripts/diff_2D_react.jl:74)."""
    spec = SCRIPTS[name]
    out = []
    if result.T_eff is not None:
        out.append("PERF: Time = %1.3f sec, T_eff = %1.2f GB/s (niter = %d)"
                   % (result.t_toc, _round_sig(result.T_eff, 3), result.niter))
    if spec.get("react"):
        if spec["ndim"] == 1:
            out.append("nx = %d, iterations tot = %d " % (n[0], result.ittot))
        else:
            out.append("nx = %d, ny = %d, iterations tot = %d " % (n[0], n[1], result.ittot))
        return out
    line = "Total time = %1.2f, time steps = %d, nx = %d, iterations tot = %d" % (
        _round_sig(spec["ttot"], 2), len(result.iters), n[0], result.ittot)
    if name == "diff_1D_lin":
        line += ", error vs analytic = %1.2e" % analytic_error_1d(result, spec)
    out.append(line + " ")
    return out


def main(argv=None):
    """CLI mirror of running a script (`python __graft_entry__.py script diff_3D_nonlin_perf` from the
    repository root) with NX/NY/NZ, DO_SAVE and DO_SAVE_VIZ from the environment
    (scripts/diff_3D/diff_3D_nonlin_perf.jl:1-8): prints the script's summary lines, appends its row to
    output/<text_file_name> and writes out_visu/<viz_file_name>. PTD_ARITH=fast runs the library's FAST
    arithmetic instead of the reference's operation order (exact, the default)."""
    argv = sys.argv[1:] if argv is None else argv
    if not argv or argv[0] not in SCRIPTS:
        print("usage: scripts.py <" + "|".join(SCRIPTS) + ">", file=sys.stderr)
        return 2
    name = argv[0]
    spec = SCRIPTS[name]
    nd = spec["ndim"]
    default_n = 64 if nd == 3 else 512
    nx = int(os.environ.get("NX", default_n))
    ny = int(os.environ.get("NY", default_n))
    nz = int(os.environ.get("NZ", default_n))
    n = [nx, ny, nz][:nd]
    arith = os.environ.get("PTD_ARITH", "exact").strip().lower()
    if arith not in ("exact", "fast"):
        print("PTD_ARITH must be exact or fast", file=sys.stderr)
        return 2
    r = run_script(name, nx, ny, nz, arith=_abi.ARITH_FAST if arith == "fast" else _abi.ARITH_EXACT)
    for line in summary_lines(name, r, n):
        print(line)
    if _env_bool("DO_SAVE", False) and not spec.get("react"):
        os.makedirs("output", exist_ok=True)
        with open(os.path.join("output", text_file_name(name)), "a") as f:
            f.write(output_row(name, r, n) + "\n")
    if _env_bool("DO_SAVE_VIZ", False):
        save_viz(name, r)
    return 0


if __name__ == "__main__":
    sys.exit(main())
