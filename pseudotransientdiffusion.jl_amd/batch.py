"""Batches of independent float64 solves through libptd's ptd_batch_* entry points, and the batched
workload of many uneven problems (2D grids from {128, 192, ..., 1024}^2 and 3D grids from
{64, 96, ..., 256}^3, D drawn from lin / step / H^3, Gaussian initial conditions with random amplitude
and centre).

No reference counterpart: the reference solves one problem per script run. Each member here is one of
the reference's converging scripts (diff_2D_{lin,linstep,nonlin}, diff_3D_{lin,linstep,nonlin}_multixpu)
on its own grid, and its result equals ptd_solve on that problem alone, bit for bit.
"""
import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import check
from .core import field_shape
from .scripts import SCRIPTS, a_eff_gb, make_params, step_field


class Batch:
    """Owns a ptd_batch: n independent problems solved together on one GPU."""

    def __init__(self, params, device=0):
        self._lib = _abi.load()
        self.params = list(params)
        arr = (_abi.Params * len(self.params))(*self.params)
        h = ctypes.c_void_p()
        check(self._lib.ptd_batch_create(arr, len(self.params), device, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ptd_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.params)

    @property
    def device_bytes(self):
        return int(self._lib.ptd_batch_device_bytes(self._h))

    def set(self, member, field, array):
        fid = _field_id(field)
        shp = field_shape(self.params[member], fid)
        a = np.asfortranarray(np.asarray(array, dtype=np.float64))
        if a.shape != shp:
            raise ValueError(f"member {member} field {field}: expected shape {shp}, got {a.shape}")
        check(self._lib.ptd_batch_set_field(self._h, member, fid, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            a.size))

    def get(self, member, field):
        fid = _field_id(field)
        shp = field_shape(self.params[member], fid)
        out = np.empty(shp, dtype=np.float64, order="F")
        check(self._lib.ptd_batch_get_field(self._h, member, fid, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            out.size))
        return out

    def solve(self, nt, raise_nan=True):
        """nt physical steps of every member. Returns per-member iteration counts (n, nt), final errors
        (n, nt), latency [ms] (n), status (n) and the wall time of the call [ms]."""
        n = len(self.params)
        iters = np.zeros((n, nt), dtype=np.int64)
        errs = np.zeros((n, nt))
        lat = np.zeros(n)
        status = np.zeros(n, dtype=np.int32)
        total = ctypes.c_double()
        st = self._lib.ptd_batch_solve(self._h, int(nt), iters.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                       errs.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       status.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(total))
        if st != _abi.ERR_NAN or raise_nan:
            check(st)
        return iters, errs, lat, status, total.value

    def group_times(self):
        """Per member group of the last solve: dicts with the sweep kind ("3D", "2D two-level", "2D LDS", "2D"), the D
        case, the member count and the device time the group's stream took (ms, summed over replays)."""
        cap = 64
        ng = ctypes.c_int32()
        kind, dcase, mem = (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)(), (ctypes.c_int32 * cap)()
        ms = (ctypes.c_double * cap)()
        check(self._lib.ptd_batch_group_times(self._h, ctypes.byref(ng), kind, dcase, mem, ms, cap))
        names = {3: "3D", 22: "2D two-level", 20: "2D LDS", 2: "2D"}
        dnames = {0: "lin", 1: "step", 2: "H^3"}
        return [{"kind": names.get(kind[g], str(kind[g])), "D": dnames.get(dcase[g], str(dcase[g])),
                 "members": int(mem[g]), "ms": round(ms[g], 2)} for g in range(min(ng.value, cap))]

    def tiling(self):
        """(z chunk depth of the 3D launches, rows per 2D block, measured ms per candidate)."""
        tz, rows = ctypes.c_int32(), ctypes.c_int32()
        ms = (ctypes.c_double * 8)()
        check(self._lib.ptd_batch_tiling(self._h, ctypes.byref(tz), ctypes.byref(rows), ms, 8))
        return tz.value, rows.value, [v for v in ms[:4]]


def _field_id(field):
    from .core import FIELD_NAMES
    return FIELD_NAMES[field] if isinstance(field, str) else int(field)


# ---- the batched workload -----------------------------------------------------------------------

SIZES_2D = list(range(128, 1025, 64))   # {128, 192, ..., 1024}
SIZES_3D = list(range(64, 257, 32))     # {64, 96, ..., 256}
SCRIPT_OF = {(2, "lin"): "diff_2D_lin", (2, "step"): "diff_2D_linstep", (2, "nonlin"): "diff_2D_nonlin",
             (3, "lin"): "diff_3D_lin_multixpu", (3, "step"): "diff_3D_linstep_multixpu",
             (3, "nonlin"): "diff_3D_nonlin_multixpu"}


@dataclass
class Member:
    script: str
    n: int
    amp: float
    centre: tuple


def workload(count=96, seed=0, sizes_2d=SIZES_2D, sizes_3d=SIZES_3D):
    """`count` members drawn with numpy's PCG64(seed): ndim 2 or 3 with equal probability, a grid size
    uniform over the size list, D type uniform over lin / step / H^3, amplitude U(0.5, 2) and centre
    U(0.3, 0.7) * L per axis."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        nd = int(rng.integers(2, 4))
        n = int(rng.choice(sizes_2d if nd == 2 else sizes_3d))
        kind = ("lin", "step", "nonlin")[int(rng.integers(0, 3))]
        amp = float(rng.uniform(0.5, 2.0))
        centre = tuple(float(rng.uniform(0.3, 0.7)) for _ in range(nd))
        out.append(Member(SCRIPT_OF[(nd, kind)], n, amp, centre))
    return out


def member_params(m, arith=_abi.ARITH_FAST):
    spec = SCRIPTS[m.script]
    nd = spec["ndim"]
    return make_params(spec, m.n, m.n, m.n if nd == 3 else 1, arith=arith)


def member_fields(m):
    """(H0, D or None) of a member: H0 = amp * exp(-|x - c L|^2) on the cell centres (the scripts' Gaussian,
    its boundary values kept as Dirichlet values), D the script's step field for the step case."""
    spec = SCRIPTS[m.script]
    nd = spec["ndim"]
    ls = [spec["lx"], spec.get("ly", 0.0), spec.get("lz", 0.0)][:nd]
    n = m.n
    axes = []
    for d in range(nd):
        dx = ls[d] / n
        xc = dx / 2 + dx * np.arange(n)
        axes.append((xc - m.centre[d] * ls[d]) ** 2)
    r2 = axes[0][:, None] + axes[1][None, :] if nd == 2 else \
        axes[0][:, None, None] + axes[1][None, :, None] + axes[2][None, None, :]
    H0 = np.asfortranarray(m.amp * np.exp(-r2))
    D = None
    if spec["dcase"] == _abi.STEP:
        D = step_field(spec, n, n, n if nd == 3 else 1)
    return H0, D


def member_a_eff_gb(m):
    nd = SCRIPTS[m.script]["ndim"]
    return a_eff_gb(nd, [m.n] * nd)


def percentile(values, q):
    v = np.sort(np.asarray(values, dtype=float))
    if v.size == 0:
        return math.nan
    k = (v.size - 1) * q / 100.0
    lo, hi = int(math.floor(k)), int(math.ceil(k))
    return float(v[lo] + (v[hi] - v[lo]) * (k - lo))
