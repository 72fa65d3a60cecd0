"""Context: one libptd problem on one GPU (device fields + the PT loop), as a Python object.

Arrays cross the boundary in the reference layout: numpy float64, column-major, with the reference's
staggered shapes (scripts/diff_3D/diff_3D_nonlin_multixpu.jl:77-83):
    H, Hold, D : (nx[, ny[, nz]])
    qx, qx2    : (nx-1[, ny-2[, nz-2]])   qy : (nx-2, ny-1[, nz-2])   qz : (nx-2, ny-2, nz-1)
    ResH, dtau : (nx-2[, ny-2[, nz-2]])
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import check

FIELD_NAMES = {
    "H": _abi.F_H, "Hold": _abi.F_HOLD, "D": _abi.F_D,
    "qHx": _abi.F_QX, "qHy": _abi.F_QY, "qHz": _abi.F_QZ,
    "qHx2": _abi.F_QX2, "qHy2": _abi.F_QY2, "qHz2": _abi.F_QZ2,
    "ResH": _abi.F_RESH, "dtau_rho": _abi.F_DTAU,
}


def field_shape(params, field):
    nd = params.ndim
    n = [params.nx, params.ny, params.nz][:nd]
    inner = [v - 2 for v in n]
    if field in (_abi.F_H, _abi.F_HOLD, _abi.F_D):
        return tuple(n)
    if field in (_abi.F_RESH, _abi.F_DTAU):
        return tuple(inner)
    axis = (field - _abi.F_QX) % 3
    if axis >= nd:
        return None
    shp = list(inner)
    shp[axis] = n[axis] - 1
    return tuple(shp)


class Context:
    """Owns a ptd_ctx (device fields for one problem) and exposes the C ABI as methods."""

    def __init__(self, params, device=0):
        self._lib = _abi.load()
        self.params = params
        h = ctypes.c_void_p()
        check(self._lib.ptd_ctx_create(ctypes.byref(params), device, ctypes.byref(h)))
        self._h = h

    # lifetime -------------------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.ptd_ctx_destroy(self._h)
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

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("context is closed")
        return self._h

    @property
    def device_bytes(self):
        return int(self._lib.ptd_ctx_device_bytes(self.handle))

    @property
    def outer_iterations(self):
        """Float64 defect-correction iterations of the last PT solve (PTD_MIXED)."""
        return int(self._lib.ptd_ctx_outer_iterations(self.handle))

    @property
    def stream(self):
        return self._lib.ptd_ctx_stream(self.handle)

    def set_params(self, params):
        check(self._lib.ptd_ctx_set_params(self.handle, ctypes.byref(params)))
        self.params = params

    # fields ---------------------------------------------------------------------------------
    @staticmethod
    def _fid(field):
        return FIELD_NAMES[field] if isinstance(field, str) else int(field)

    def set(self, field, array):
        fid = self._fid(field)
        a = np.asarray(array, dtype=np.float64)
        shp = field_shape(self.params, fid)
        if shp is None or a.shape != shp:
            raise ValueError(f"field {field}: expected shape {shp}, got {a.shape}")
        a = np.asfortranarray(a)
        check(self._lib.ptd_set_field(self.handle, fid, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      a.size))

    def get(self, field):
        fid = self._fid(field)
        shp = field_shape(self.params, fid)
        if shp is None:
            raise ValueError(f"field {field} is absent in {self.params.ndim}D")
        out = np.empty(shp, dtype=np.float64, order="F")
        check(self._lib.ptd_get_field(self.handle, fid, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      out.size))
        return out

    # kernel-level entry points ----------------------------------------------------------------
    def compute_iter_params(self):
        check(self._lib.ptd_compute_iter_params(self.handle))

    def compute_flux(self):
        check(self._lib.ptd_compute_flux(self.handle))

    def compute_update(self):
        check(self._lib.ptd_compute_update(self.handle))

    def compute_flux_res(self):
        check(self._lib.ptd_compute_flux_res(self.handle))

    def check_res(self):
        rms, mx = ctypes.c_double(), ctypes.c_double()
        check(self._lib.ptd_check_res(self.handle, ctypes.byref(rms), ctypes.byref(mx)))
        return rms.value, mx.value

    def assign(self):
        check(self._lib.ptd_assign(self.handle))

    # loop-level entry points -------------------------------------------------------------------
    def iterate(self, niter):
        check(self._lib.ptd_iterate(self.handle, int(niter)))

    def iterate_timed(self, niter):
        ms = ctypes.c_double()
        check(self._lib.ptd_iterate_timed(self.handle, int(niter), ctypes.byref(ms)))
        return ms.value

    def iterations_per_launch(self):
        """PT iterations per launch of this context's unchecked iterations (3 three-level, 2 two-level, 1 single
        sweeps, up to 10 LDS-resident 2D)."""
        n = ctypes.c_int32()
        check(self._lib.ptd_ctx_iterations_per_launch(self.handle, ctypes.byref(n)))
        return n.value

    def pt_solve(self, raise_nan=True):
        it, rms, mx = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
        st = self._lib.ptd_pt_solve(self.handle, ctypes.byref(it), ctypes.byref(rms), ctypes.byref(mx))
        if st != _abi.ERR_NAN or raise_nan:
            check(st)
        return it.value, rms.value, mx.value

    def solve(self, nt):
        iters = np.zeros(nt, dtype=np.int64)
        errs = np.zeros(nt)
        check(self._lib.ptd_solve(self.handle, int(nt), iters.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                  errs.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return iters, errs

    def grad_solve(self, nt, h_obs):
        """ptd_grad_solve: solve nt steps and return the loss L = 1/2 sum_inner (H - h_obs)^2, dL/dD (None for
        D = H^3) and dL/dH0 (reference layout) and the forward / adjoint PT iteration counts."""
        shp = field_shape(self.params, _abi.F_H)
        hobs = np.asfortranarray(np.asarray(h_obs, dtype=np.float64))
        if hobs.shape != shp:
            raise ValueError(f"h_obs: expected shape {shp}, got {hobs.shape}")
        loss = ctypes.c_double()
        gD = None if self.params.dcase == _abi.NONLIN else np.zeros(shp, order="F")
        gH0 = np.zeros(shp, order="F")
        itf = np.zeros(nt, dtype=np.int64)
        ita = np.zeros(nt, dtype=np.int64)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int64)
        check(self._lib.ptd_grad_solve(self.handle, int(nt), hobs.ctypes.data_as(dp), hobs.size, ctypes.byref(loss),
                                       None if gD is None else gD.ctypes.data_as(dp), gH0.ctypes.data_as(dp),
                                       itf.ctypes.data_as(ip),
                                       ita.ctypes.data_as(ip)))
        return {"loss": loss.value, "grad_D": gD, "grad_H0": gH0, "iters_fwd": itf, "iters_adj": ita}
