"""ctypes front end of the CPU oracle (oracle/libptd_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg,
never by the product package.

SCRIPTS restates the parameter blocks of the reference scripts (file:line cited per entry); run_script()
reproduces what `include("scripts/<name>.jl")` returns with USE_RETURN=true: the cell-centre x
coordinates and the final H (column-major, shape (nx[,ny[,nz]])).
"""
import ctypes
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libptd_oracle.so")

LIN, STEP, NONLIN = 0, 1, 2
PLAIN, PERF = 0, 1


class Case(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int32), ("dcase", ctypes.c_int32), ("form", ctypes.c_int32),
        ("nout", ctypes.c_int32),
        ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
        ("itmax", ctypes.c_int64), ("nt", ctypes.c_int32), ("pad_", ctypes.c_int32),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("dt", ctypes.c_double), ("tol", ctypes.c_double), ("cfl", ctypes.c_double),
        ("resc", ctypes.c_double), ("d0", ctypes.c_double), ("dmax", ctypes.c_double),
        ("norm_len", ctypes.c_double),
    ]


class Derived(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "dx", "dy", "dz", "_dx", "_dy", "_dz", "_dt", "vpdt", "max_l", "max_l2",
        "re", "theta_r", "dtau_rho")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        cp = ctypes.POINTER(Case)
        L.ptdo_derive.argtypes = [cp, ctypes.POINTER(Derived)]
        L.ptdo_linrange.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int64, dp]
        L.ptdo_ic.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                              ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                              ctypes.c_double, dp]
        L.ptdo_dstep.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, dp]
        L.ptdo_iter_params.argtypes = [cp, dp, dp]
        L.ptdo_flux.argtypes = [cp, dp, dp, dp, dp, dp, dp, dp, dp]
        L.ptdo_update.argtypes = [cp, dp, dp, dp, dp, dp, dp]
        L.ptdo_flux_res.argtypes = [cp, dp, dp, dp, dp, dp]
        L.ptdo_check_res.argtypes = [cp, dp, dp, dp, dp, dp, dp, dp]
        L.ptdo_check_res.restype = ctypes.c_double
        L.ptdo_solve.argtypes = [cp, dp, dp, dp, ctypes.POINTER(ctypes.c_int64), dp]
        L.ptdo_solve.restype = ctypes.c_int
        L.ptdo_iterate.argtypes = [cp, dp, dp, dp, dp, dp, dp, dp, ctypes.c_int64]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["F_CONTIGUOUS"]:
        raise ValueError("oracle arrays must be float64, column-major")
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# Parameter blocks of the reference scripts. ic: 0 shifted LinRange (1D/2D), 1 xc comprehension (3D perf),
# 2 x_g (3D multixpu); dstep: 0 (1D/2D index rule), 1 (3D x_g rule).
SCRIPTS = {
    # scripts/diff_1D_lin.jl:34-50
    "diff_1D_lin": dict(ndim=1, dcase=LIN, form=PLAIN, lx=10.0, d0=1.0, dt=0.2, ttot=1.0, tol=1e-8,
                        itmax=100000, nout=10, cfl=0.99, resc=1.0, ic=0),
    # scripts/diff_1D_linstep.jl:45-61
    "diff_1D_linstep": dict(ndim=1, dcase=STEP, form=PLAIN, lx=10.0, d1=1.0, d2=1e-4, dt=0.2, ttot=1.0,
                            tol=1e-8, itmax=100000, nout=10, cfl=0.99, resc=1.0, ic=0, dstep=0),
    # scripts/diff_1D_nonlin.jl:40-53 (CFL 1.0, no Resc)
    "diff_1D_nonlin": dict(ndim=1, dcase=NONLIN, form=PLAIN, lx=10.0, dt=0.2, ttot=1.0, tol=1e-8,
                           itmax=100000, nout=10, cfl=1.0, resc=1.0, ic=0),
    # scripts/diff_1D_react.jl:38-53: lin with dt -> theta_kin, Hold -> Heq = H0, H starts at 0, one solve
    "diff_1D_react": dict(ndim=1, dcase=LIN, form=PLAIN, lx=20.0, d0=1.0, dt=0.1, nt=1, tol=1e-8,
                          itmax=100000, nout=10, cfl=0.99, resc=1.0, ic=0, react=True),
    # scripts/diff_2D_lin.jl:38-53
    "diff_2D_lin": dict(ndim=2, dcase=LIN, form=PLAIN, lx=10.0, ly=10.0, d0=1.0, dt=0.2, ttot=1.0,
                        tol=1e-8, itmax=100000, nout=10, cfl=1 / math.sqrt(2), resc=1.0, ic=0),
    # scripts/diff_2D_linstep.jl:55-83
    "diff_2D_linstep": dict(ndim=2, dcase=STEP, form=PLAIN, lx=10.0, ly=10.0, d1=1.0, d2=1e-4, dt=0.2,
                            ttot=1.0, tol=1e-8, itmax=100000, nout=10, cfl=1 / math.sqrt(2), resc=1.0,
                            ic=0, dstep=0),
    # scripts/diff_2D_nonlin.jl:46-58
    "diff_2D_nonlin": dict(ndim=2, dcase=NONLIN, form=PLAIN, lx=10.0, ly=10.0, dt=0.2, ttot=1.0, tol=1e-8,
                           itmax=100000, nout=10, cfl=1 / math.sqrt(2), resc=1 / 1.2, ic=0),
    # scripts/diff_2D_nonlin_perf.jl:56-71 (itMax = 1e2, nout = 2000: fixed 2 x 100 iterations)
    "diff_2D_nonlin_perf": dict(ndim=2, dcase=NONLIN, form=PERF, lx=10.0, ly=10.0, dt=0.2, ttot=0.4,
                                tol=1e-8, itmax=100, nout=2000, cfl=1 / math.sqrt(2), resc=1 / 1.2, ic=0),
    # scripts/diff_2D_react.jl:38-53
    "diff_2D_react": dict(ndim=2, dcase=LIN, form=PLAIN, lx=10.0, ly=10.0, d0=1.0, dt=0.1, nt=1, tol=1e-8,
                          itmax=100000, nout=10, cfl=1 / math.sqrt(2), resc=1.0, ic=0, react=True),
    # scripts/diff_3D/diff_3D_lin_multixpu.jl:45-62
    "diff_3D_lin_multixpu": dict(ndim=3, dcase=LIN, form=PLAIN, lx=10.0, ly=10.0, lz=10.0, d0=1.0, dt=0.2,
                                 ttot=1.0, tol=1e-8, itmax=100000, nout=10, cfl=1 / math.sqrt(3),
                                 resc=1.0, ic=2),
    # scripts/diff_3D/diff_3D_linstep_multixpu.jl:67-104
    "diff_3D_linstep_multixpu": dict(ndim=3, dcase=STEP, form=PLAIN, lx=10.0, ly=10.0, lz=10.0, d1=1.0,
                                     d2=1e-4, dt=0.2, ttot=1.0, tol=1e-8, itmax=100000, nout=10,
                                     cfl=1 / math.sqrt(3), resc=1.0, ic=2, dstep=1),
    # scripts/diff_3D/diff_3D_nonlin_multixpu.jl:57-75
    "diff_3D_nonlin_multixpu": dict(ndim=3, dcase=NONLIN, form=PLAIN, lx=10.0, ly=10.0, lz=10.0, dt=0.2,
                                    ttot=1.0, tol=1e-8, itmax=100000, nout=10, cfl=1 / math.sqrt(3),
                                    resc=1 / 1.2, ic=2),
    # scripts/diff_3D/diff_3D_nonlin_perf.jl:61-78 (fixed 2 x 1000 iterations)
    "diff_3D_nonlin_perf": dict(ndim=3, dcase=NONLIN, form=PERF, lx=10.0, ly=10.0, lz=10.0, dt=0.2,
                                ttot=0.4, tol=1e-8, itmax=1000, nout=2000, cfl=1 / math.sqrt(3),
                                resc=1 / 1.2, ic=1),
    # scripts/diff_3D/diff_3D_nonlin_multixpu_perf.jl:61-80
    "diff_3D_nonlin_multixpu_perf": dict(ndim=3, dcase=NONLIN, form=PERF, lx=10.0, ly=10.0, lz=10.0,
                                         dt=0.2, ttot=0.4, tol=1e-8, itmax=1000, nout=2000,
                                         cfl=1 / math.sqrt(3), resc=1 / 1.2, ic=2),
}


def make_case(spec, nx, ny=1, nz=1, **over):
    s = dict(spec)
    s.update(over)
    nt = s.get("nt")
    if nt is None:
        nt = int(math.ceil(s["ttot"] / s["dt"]))  # nt = Int(ceil(ttot/dt))
    c = Case()
    c.ndim, c.dcase, c.form, c.nout = s["ndim"], s["dcase"], s["form"], s["nout"]
    c.nx, c.ny, c.nz = nx, (ny if s["ndim"] >= 2 else 1), (nz if s["ndim"] >= 3 else 1)
    c.itmax, c.nt = int(s["itmax"]), nt
    c.lx, c.ly, c.lz = s["lx"], s.get("ly", 0.0), s.get("lz", 0.0)
    c.dt, c.tol, c.cfl, c.resc = s["dt"], s["tol"], s["cfl"], s["resc"]
    c.d0 = s.get("d0", 1.0)
    c.dmax = max(s.get("d1", 1.0), s.get("d2", 1.0))
    c.norm_len = s.get("norm_len", 0.0)
    return c


def derive(case):
    d = Derived()
    lib().ptdo_derive(ctypes.byref(case), ctypes.byref(d))
    return d


def shape_of(case):
    return tuple(int(n) for n in (case.nx, case.ny, case.nz)[:case.ndim])


def linrange(a, b, n):
    out = np.empty(n)
    lib().ptdo_linrange(a, b, n, _p(out))
    return out


def initial_condition(kind, case, amp=2.0, bval=1.0):
    H0 = np.zeros(shape_of(case), order="F")
    lib().ptdo_ic(kind, case.ndim, case.nx, case.ny, case.nz, case.lx, case.ly, case.lz, amp, bval, _p(H0))
    return H0


def step_field(kind, case, d1=1.0, d2=1e-4):
    D = np.zeros(shape_of(case), order="F")
    lib().ptdo_dstep(kind, case.ndim, case.nx, case.ny, case.nz, case.lx, case.ly, case.lz, d1, d2, _p(D))
    return D


def solve(case, H, Hold, D=None):
    """Run the script loop in place on H/Hold; returns (iters per step, final err per step, status)."""
    iters = np.zeros(case.nt, dtype=np.int64)
    errs = np.zeros(case.nt)
    st = lib().ptdo_solve(ctypes.byref(case), _p(H), _p(Hold), _p(D),
                          iters.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _p(errs))
    return iters, errs, st


def run_script(name, nx, ny=1, nz=1, **over):
    """Equivalent of include("scripts/<name>.jl") with USE_RETURN=true and NX/NY/NZ set."""
    spec = dict(SCRIPTS[name])
    spec.update(over)
    case = make_case(spec, nx, ny, nz)
    d = derive(case)
    xc = linrange(d.dx / 2, case.lx - d.dx / 2, case.nx)
    H0 = initial_condition(spec["ic"], case)
    D = step_field(spec["dstep"], case, spec["d1"], spec["d2"]) if spec["dcase"] == STEP else None
    if spec.get("react"):
        H = np.zeros_like(H0)      # H = @zeros(nx): boundaries stay 0 (diff_1D_react.jl:52)
        Hold = H0.copy(order="F")  # Heq = H0
    else:
        H = H0.copy(order="F")
        Hold = H0.copy(order="F")
    iters, errs, st = solve(case, H, Hold, D)
    return dict(xc=xc, H=H, H0=H0, D=D, iters=iters, errs=errs, status=st, case=case)


def sample_indices(n):
    """inds = Int.(ceil.(LinRange(1, n, 12))) (test/test_1D.jl:23), returned 0-based."""
    r = linrange(1.0, float(n), 12)
    return np.array([int(math.ceil(v)) - 1 for v in r])
