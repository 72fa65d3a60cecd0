"""MI355X-native accelerated pseudo-transient diffusion (libptd) -- Python host side.

The compute path is libptd.so (hand-written gfx950 HIP kernels behind the C ABI in include/ptd.h).
This package binds it with ctypes and restates the reference scripts' host code (parameters, initial
conditions, the physical-time loop) on top of it.
"""
from . import _abi
from ._abi import (FP32, FP64, LIN, MIXED, NONLIN, NORM_MAX, NORM_RMS, PERF, PLAIN, STEP, Params, PTDError,
                   load)
from .batch import Batch
from .core import Context, field_shape
from .scripts import SCRIPTS, ScriptResult, a_eff_gb, initial_condition, make_params, run_script, step_field

__all__ = [
    "Batch", "Context", "Params", "PTDError", "SCRIPTS", "ScriptResult", "a_eff_gb", "field_shape",
    "initial_condition", "load", "make_params", "run_script", "step_field",
    "FP32", "FP64", "MIXED", "LIN", "STEP", "NONLIN", "PLAIN", "PERF", "NORM_RMS", "NORM_MAX",
]
