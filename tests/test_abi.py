"""The C-ABI boundary without a GPU: exports, struct layouts, host-only calls.

- every function declared in include/sle_api.h and include/sle_ext.h is
  exported by both libsle.so (SLE_Real = float) and libsle_f64.so (double);
- sizeof/offsetof of every ABI struct, as a C compiler lays them out from the
  header, match the ctypes mirror used by the tests (both SLE_Real widths);
- model building, id lookups, config file I/O and error statuses work on the
  host; sle_Initialize reports SLE_ERROR_CUDA_FAILURE when no GPU is visible.
"""
import ctypes as C
import os
import re
import subprocess

import pytest

from state_estimation_with_cuda_amd import gridgen as gg
from state_estimation_with_cuda_amd import sle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("sle_api.h", "sle_ext.h")]


def declared():
    names = set()
    for h in HEADERS:
        txt = open(h).read()
        names |= set(re.findall(r"SLE_CALL\s+(sle_\w+)\s*\(", txt))
    return names


@pytest.mark.parametrize("lib", ["libsle.so", "libsle_f64.so"])
def test_exports_match_headers(lib):
    path = os.path.join(sle.LIB_DIR, lib)
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.split() and ln.split()[-1].startswith("sle_")}
    decl = declared()
    assert decl == exported
    assert set(sle.API_FUNCTIONS) | set(sle.EXT_FUNCTIONS) == decl
    assert len(sle.API_FUNCTIONS) == 74


STRUCTS = ["BusInfo", "BranchInfo", "MeasurementInfo", "SwitchInfo", "MeterInfo", "MeterReading", "Result",
           "Config", "SolverConfig", "MeasurementUpdate", "MeasurementDetails"]


@pytest.mark.parametrize("double", [False, True])
def test_struct_layout_matches_c_compiler(tmp_path, double):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sle_api.h"', "int main(void){"]
    L = sle.library(double)
    for name in STRUCTS:
        st = L.S[name]
        lines.append(f'printf("{name} %zu\\n", sizeof(SLE_{name}));')
        for field, _ in st._fields_:
            lines.append(f'printf("{name}.{field} %zu\\n", offsetof(SLE_{name}, {field}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    cmd = ["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)]
    if double:
        cmd.insert(1, "-DSLE_USE_DOUBLE")
    subprocess.run(cmd, check=True)
    got = dict(ln.rsplit(" ", 1) for ln in subprocess.run([str(exe)], capture_output=True, text=True,
                                                          check=True).stdout.splitlines())
    for name in STRUCTS:
        st = L.S[name]
        assert int(got[name]) == C.sizeof(st), name
        for field, _ in st._fields_:
            assert int(got[f"{name}.{field}"]) == getattr(st, field).offset, (name, field)


def test_host_only_model_building_and_errors():
    e = sle.Engine(double=True, initialize=False)
    L, S = e.L, e.S
    assert L.sle_GetVersion() == b"0.20.0"
    assert L.sle_LoadIEEE14Bus(e.h) == 0
    assert L.sle_GetBusCount(e.h) == 14 and L.sle_GetBranchCount(e.h) == 20
    assert L.sle_GetMeasurementCount(e.h) == 31
    bus = S["BusInfo"](b"Bus1", 69.0, 0, 1.0, 0.0, 0.0)
    assert L.sle_AddBus(e.h, C.byref(bus)) == -12  # duplicate id
    br = S["BranchInfo"](b"Lx", b"Bus1", b"NoSuchBus", 0.01, 0.1, 0.0, 0.0, 1.0, 0.0, 0, None)
    assert L.sle_AddBranch(e.h, C.byref(br)) == -11  # element not found
    assert L.sle_AddBus(None, C.byref(bus)) == -4
    assert abs(L.sle_GetTransformerTap(e.h, b"Xfmr4-7") - 0.978) < 1e-12
    assert L.sle_SetTransformerTap(e.h, b"Xfmr4-7", 0.99) == 0
    assert abs(L.sle_GetTransformerTap(e.h, b"Xfmr4-7") - 0.99) < 1e-12
    assert L.sle_GetTransformerTap(e.h, b"missing") == 1.0
    assert L.sle_GetMeasurementIdByIndex(e.h, 14) == b"P1"
    sw = S["SwitchInfo"](b"SW1", b"Line1-2", 1)
    assert L.sle_AddSwitch(e.h, C.byref(sw)) == 0
    assert L.sle_GetSwitchStatus(e.h, b"SW1") == 1
    assert L.sle_SetSwitchStatus(e.h, b"SW1", 0) == 0
    assert L.sle_GetSwitchStatus(e.h, b"SW1") == 0
    if not sle.gpu_available():
        assert L.sle_Initialize(e.h) == -3
        assert b"HIP" in L.sle_GetLastError(e.h)
        res = S["Result"]()
        assert L.sle_Solve(e.h, 1, C.byref(res)) == -2
    e.close()
    sle.library(True).L.sle_Destroy(None)  # NULL is a no-op


def test_config_file_roundtrip(tmp_path):
    L = sle.library(True)
    cfg = sle.Engine(double=True, initialize=False).default_config()
    cfg.huber_gamma = 2.25
    cfg.max_realtime_iterations = 7
    cfg.use_double_precision = 1
    p = str(tmp_path / "x.config").encode()
    assert L.L.sle_SaveConfigFile(p, C.byref(cfg)) == 0
    back = L.S["Config"]()
    assert L.L.sle_LoadConfigFile(p, C.byref(back)) == 0
    assert back.huber_gamma == 2.25 and back.max_realtime_iterations == 7 and back.use_double_precision == 1
    assert L.L.sle_LoadConfigFile(b"/nonexistent/file", C.byref(back)) == -13
    h = C.c_void_p()
    assert L.L.sle_CreateFromConfigFile(C.byref(h), p) == 0
    got = L.S["Config"]()
    assert L.L.sle_GetConfig(h, C.byref(got)) == 0 and got.huber_gamma == 2.25
    L.L.sle_Destroy(h)


def test_float_abi_bulk_load_matches_double():
    g = gg.grid14(0)
    for double in (False, True):
        e = sle.Engine(double=double, initialize=False)
        e.load_grid(g)
        assert e.L.sle_GetMeasurementCount(e.h) == g.m
        d = e.S["MeasurementDetails"]()
        assert e.L.sle_GetMeasurementDetails(e.h, b"M30", C.byref(d)) == 0
        assert d.type == g.m_type[30] and d.internal_index == 30
        assert abs(d.sigma - g.sigma[30]) < 1e-7
        e.close()


def build_lifecycle(tmp_path, double=False):
    exe = tmp_path / ("lifecycle_f64" if double else "lifecycle")
    lib = "sle_f64" if double else "sle"
    cmd = ["gcc", "-std=c99", "-O1", "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "c", "lifecycle.c"),
           "-L", sle.LIB_DIR, f"-l{lib}", f"-Wl,-rpath,{sle.LIB_DIR}", "-o", str(exe)]
    if double:
        cmd.insert(1, "-DSLE_USE_DOUBLE")
    subprocess.run(cmd, check=True)
    return exe


@pytest.mark.parametrize("double", [False, True])
def test_c_host_links_and_runs_host_side(tmp_path, double):
    """A plain C program, compiled against include/sle_api.h and linked with the
    library, runs the reference's lifecycle; without a GPU it stops at
    sle_Initialize with the documented status."""
    exe = build_lifecycle(tmp_path, double)
    if sle.gpu_available():
        pytest.skip("covered by the GPU run of the same program")
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "nogpu"


def test_options_documented_in_header_are_accepted():
    """Every sle_ext_SetOption key listed in include/sle_ext.h is accepted before
    a device exists; unknown keys and out-of-range values are rejected."""
    import re

    text = open(os.path.join(ROOT, "include", "sle_ext.h")).read()
    block = text[: text.index("sle_ext_SetOption")]
    keys = sorted(set(re.findall(r'^ \*   "([a-z_0-9]+)"', block, re.M)))
    assert {"pcg_steps", "pinv", "pinv_dense", "lds_flags", "xcd_remap", "meas_by_type"} <= set(keys), keys
    e = sle.Engine(double=True, initialize=False)
    for k in keys:
        e.set_option(k, 1)
    with pytest.raises(sle.SLEError):
        e.set_option("no_such_option", 1)
    with pytest.raises(sle.SLEError):
        e.set_option("pcg_steps", 99)
    with pytest.raises(sle.SLEError):
        e.set_option("pinv_wg_fwd", 3)  # waves per middle subtree: 0, 1, 4, 8 or 16
    with pytest.raises(sle.SLEError):
        e.set_option("queue_plain", 3)  # 0, 1 or 2
    with pytest.raises(sle.SLEError):
        e.set_option("fork_sub_m", 3)  # 0, 1 or 2
    e.close()
