"""The boundary against the reference's own header, without a GPU.

tests/golden/abi_reference.json holds what tests/golden/abi_probe.py read from
the reference's include/sle_api.h (v0.20.0): the 74 function signatures
(parameter types), every enumerator's value, and sizeof / offsetof of every
struct for SLE_Real = float and = double, as gcc lays them out.  This
repository's include/sle_api.h must give the same facts, both libraries must
export every one of the 74 functions, and the ctypes mirror the tests use must
lay the structs out the same way.
"""
import ctypes as C
import json
import os
import subprocess
import sys

import pytest

from state_estimation_with_cuda_amd import sle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, GOLDEN)
import abi_probe  # noqa: E402

REF = json.load(open(os.path.join(GOLDEN, "abi_reference.json")))
OURS = abi_probe.probe(os.path.join(ROOT, "include", "sle_api.h"))


def test_fixture_is_the_full_reference_abi():
    assert len(REF["functions"]) == 74
    assert len(REF["enums"]) == 43
    assert len(REF["structs"]) == 11


def test_function_signatures_match_the_reference():
    assert set(OURS["functions"]) == set(REF["functions"])
    for name, sig in REF["functions"].items():
        assert OURS["functions"][name] == sig, name


def test_enumerators_and_struct_fields_match_the_reference():
    assert OURS["enums"] == REF["enums"]
    assert OURS["structs"] == REF["structs"]


@pytest.mark.parametrize("width", ["float", "double"])
def test_layouts_and_enum_values_match_the_reference(width):
    assert OURS["layouts"][width] == REF["layouts"][width]


@pytest.mark.parametrize("lib", ["libsle.so", "libsle_f64.so"])
def test_libraries_export_every_reference_function(lib):
    path = os.path.join(sle.LIB_DIR, lib)
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.split()}
    missing = set(REF["functions"]) - exported
    assert not missing, sorted(missing)


@pytest.mark.parametrize("double", [False, True])
def test_ctypes_mirror_matches_the_reference_layout(double):
    lay = REF["layouts"]["double" if double else "float"]
    L = sle.library(double)
    for name, fields in REF["structs"].items():
        st = L.S[name[len("SLE_"):]]
        assert C.sizeof(st) == lay[f"sizeof {name}"], name
        assert [f for f, _ in st._fields_] == fields, name
        for f in fields:
            assert getattr(st, f).offset == lay[f"offsetof {name}.{f}"], (name, f)
