"""The C-ABI library loads, exports every function include/ptd.h declares, and validates arguments
before touching a device (no GPU needed)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "ptd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ptd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function(pkg):
    lib = pkg.load()
    declared = _declared_functions()
    assert len(declared) >= 20
    from importlib import import_module
    abi = import_module("ptd_amd._abi")
    assert sorted(abi.EXPORTS) == declared
    for name in declared:
        assert hasattr(lib, name), f"{name} not exported by libptd.so"


def test_abi_version(pkg):
    assert pkg.load().ptd_abi_version() == 1


def test_struct_layout_matches_header(pkg):
    # ptd_params: 4 int32, 4 int64, 4 int32, 14 doubles
    assert ctypes.sizeof(pkg.Params) == 4 * 4 + 4 * 8 + 4 * 4 + 14 * 8


@pytest.mark.parametrize("field,value", [("ndim", 4), ("dcase", 7), ("form", 3), ("precision", 9),
                                         ("nx", 2), ("nout", 0), ("dx", 0.0), ("itmax", -1),
                                         ("arith", 5)])
def test_invalid_params_rejected_before_device(field, value, pkg):
    from importlib import import_module
    abi = import_module("ptd_amd._abi")
    p = pkg.make_params(pkg.SCRIPTS["diff_3D_nonlin_perf"], 16, 16, 16)
    setattr(p, field, value)
    with pytest.raises(pkg.PTDError) as e:
        pkg.Context(p)
    assert e.value.status == abi.ERR_ARG


def test_null_context_is_an_argument_error(pkg):
    lib = pkg.load()
    from importlib import import_module
    abi = import_module("ptd_amd._abi")
    assert lib.ptd_iterate(None, 1) == abi.ERR_ARG
    assert lib.ptd_check_res(None, None, None) == abi.ERR_ARG
    assert b"NULL" in lib.ptd_last_error()


@pytest.mark.parametrize("what", ["1d", "fp32", "perf", "odd_nout", "itmax", "nout_mismatch", "empty",
                                  "big_plane"])
def test_batch_rejects_invalid_members_before_device(what, pkg):
    from importlib import import_module
    abi = import_module("ptd_amd._abi")
    lin2 = pkg.make_params(pkg.SCRIPTS["diff_2D_lin"], 32, 32)
    nl3 = pkg.make_params(pkg.SCRIPTS["diff_3D_nonlin_multixpu"], 16, 16, 16)
    members = [lin2, nl3]
    if what == "1d":
        members[0] = pkg.make_params(pkg.SCRIPTS["diff_1D_lin"], 32)
    elif what == "fp32":
        members[1].precision = abi.FP32
    elif what == "perf":
        members[1].form = abi.PERF
    elif what == "odd_nout":
        for m in members:
            m.nout = 5
    elif what == "itmax":
        members[0].itmax = 1005
    elif what == "nout_mismatch":
        members[1].nout = 20
    elif what == "empty":
        members = []
    elif what == "big_plane":
        # 37 x-y planes of 4000^2 float64 exceed the 4 GiB span of the batched sweep's 32-bit offsets
        members[1] = pkg.make_params(pkg.SCRIPTS["diff_3D_nonlin_multixpu"], 4000, 4000, 16)
    with pytest.raises(pkg.PTDError) as e:
        pkg.Batch(members)
    assert e.value.status == abi.ERR_ARG
    if what == "big_plane":
        assert b"plane too large" in pkg.load().ptd_last_error()


@pytest.mark.parametrize("nxy,ok", [(511, True), (8000, True), (9500, False), (30000, False)])
def test_mixed_rejects_planes_beyond_the_sweep_offsets(nxy, ok, pkg):
    """PTD_MIXED always runs the fused 3D correction sweeps, which address (tz + 5) x-y planes from a chunk base
    with 32-bit byte offsets: a float64 plane of nx*ny cells must satisfy 6 * nx*ny * 8 < 2^32."""
    from importlib import import_module
    abi = import_module("ptd_amd._abi")
    assert ok == (6 * nxy * nxy * 8 < 2 ** 32)
    p = pkg.make_params(pkg.SCRIPTS["diff_3D_nonlin_multixpu"], nxy, nxy, 4, precision=abi.MIXED)
    if ok:
        return  # creating it would allocate; the rule above is what the library applies
    with pytest.raises(pkg.PTDError) as e:
        pkg.Context(p)
    assert e.value.status == abi.ERR_ARG
    assert b"4 GiB" in pkg.load().ptd_last_error()


JULIA_DIR = os.path.join(ROOT, "pseudotransientdiffusion.jl_amd", "julia")


def test_julia_binding_matches_header(pkg):
    """The Julia binding (LibPTD.jl; Julia is not installed here, so this is a static check): every ccall
    names a function include/ptd.h declares, every declared function but the device-buffer entry point is
    bound, PtdParams has ptd_params' fields in order with the same widths, and the enum values agree."""
    from importlib import import_module
    abi = import_module("ptd_amd._abi")
    src = open(os.path.join(JULIA_DIR, "LibPTD.jl")).read()
    called = set(re.findall(r"ccall\(\(:(ptd_[a-z0-9_]+),\s*libptd\)", src))
    declared = set(_declared_functions())
    assert called <= declared, called - declared
    assert declared - called == {"ptd_pt_solve_noalloc"}, declared - called
    body = re.search(r"struct PtdParams\n(.*?)\nend", src, flags=re.S).group(1)
    fields = re.findall(r"^\s+(\w+)::(\w+)\s*$", body, flags=re.M)
    jl_width = {"Int32": 4, "Int64": 8, "Float64": 8}
    assert [(n, jl_width[t]) for n, t in fields] == [(n, ctypes.sizeof(t)) for n, t in abi.Params._fields_]
    header = open(os.path.join(ROOT, "include", "ptd.h")).read()
    consts = {}
    for names, values in re.findall(r"^const ([A-Z0-9_, ]+) = ([\d, ]+)$", src, flags=re.M):
        consts.update(zip([v.strip() for v in names.split(",")], [int(v) for v in values.split(",")]))
    enums = re.findall(r"\b(PTD_[A-Z0-9_]+)\s*=\s*(\d+)", header)
    assert len(enums) >= 30
    for name, value in enums:
        assert consts.get(name) == int(value), name


def test_julia_script_keeps_the_reference_interface():
    """The drop-in script keeps the reference script's environment flags, numerics, printed line and row."""
    src = open(os.path.join(JULIA_DIR, "diff_3D_nonlin_multixpu.jl")).read()
    for flag in ("USE_RETURN", "DO_SAVE", "NX", "NY", "NZ"):
        assert f'ENV["{flag}"]' in src
    for line in ("tol        = 1e-8", "itMax      = 1e5", "nout       = 10", "CFL        = 1 / sqrt(3)",
                 "Resc       = 1 / 1.2", "dt         = 0.2", "ttot       = 1.0"):
        assert line in src
    assert '"Total time = %1.2f, time steps = %d, nx = %d, iterations tot = %d \\n"' in src
    assert "out_diff_3D_nonlin.txt" in src and "LibPTD.solve!(ctx, nt)" in src
