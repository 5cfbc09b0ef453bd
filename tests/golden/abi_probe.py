"""Extracts the C-ABI facts of an sle_api.h: function signatures, enum values and
the layout (sizeof / offsetof) of every struct, for both SLE_Real widths.

Run as a script in the build container, it reads the reference's
include/sle_api.h and writes tests/golden/abi_reference.json, the fixture
tests/test_abi_reference.py compares this repository's header and libraries
with.  The layouts come from gcc compiling a small probe program that includes
the header (the header is plain C: stdint.h / stddef.h only); signatures and
names come from the header text.
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "abi_reference.json")
REFERENCE_HEADER = "/root/reference/include/sle_api.h"


def _strip_comments(txt: str) -> str:
    """Comments and preprocessor lines removed."""
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", " ", txt)
    return "\n".join(ln for ln in txt.splitlines() if not ln.lstrip().startswith("#"))


def _param_type(p: str) -> str:
    """Parameter declaration without its name: 'const SLE_Real* values' -> 'const SLE_Real*'."""
    p = " ".join(p.replace("*", " * ").split())
    toks = p.split(" ")
    if len(toks) > 1 and re.fullmatch(r"[A-Za-z_]\w*", toks[-1]) and toks[-2] != "struct":
        toks = toks[:-1]
    return " ".join(toks).replace(" *", "*")


def functions(header_text: str) -> dict:
    txt = _strip_comments(header_text)
    out = {}
    for m in re.finditer(r"SLE_API\s+(.*?)\s+SLE_CALL\s+(sle_\w+)\s*\((.*?)\)\s*;", txt, flags=re.S):
        ret = _param_type(m.group(1) + " x")
        params = [q.strip() for q in m.group(3).split(",")]
        out[m.group(2)] = ret + "(" + ", ".join(_param_type(q) for q in params) + ")"
    return out


def enums(header_text: str) -> list:
    txt = _strip_comments(header_text)
    names = []
    for m in re.finditer(r"typedef\s+enum\s+\w*\s*\{(.*?)\}", txt, flags=re.S):
        for item in m.group(1).split(","):
            item = item.strip()
            if item:
                names.append(item.split("=")[0].strip())
    return names


def structs(header_text: str) -> dict:
    txt = _strip_comments(header_text)
    out = {}
    for m in re.finditer(r"typedef\s+struct\s+\w+\s*\{(.*?)\}\s*(\w+)\s*;", txt, flags=re.S):
        fields = []
        for decl in m.group(1).split(";"):
            decl = " ".join(decl.split())
            if decl:
                fields.append(re.findall(r"([A-Za-z_]\w*)\s*(?:\[[^\]]*\])?\s*$", decl)[0])
        out[m.group(2)] = fields
    return out


def layouts(header_path: str) -> dict:
    """sizeof / offsetof of every struct and the value of every enumerator, for
    SLE_Real = float and = double, as gcc lays them out from the header."""
    txt = open(header_path).read()
    st = structs(txt)
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "sle_api.h"', "int main(void){"]
    for name, fields in st.items():
        lines.append(f'printf("sizeof {name} %zu\\n", sizeof({name}));')
        for f in fields:
            lines.append(f'printf("offsetof {name}.{f} %zu\\n", offsetof({name}, {f}));')
    for e in enums(txt):
        lines.append(f'printf("enum {e} %ld\\n", (long){e});')
    lines.append('printf("sizeof SLE_Real %zu\\n", sizeof(SLE_Real));')
    lines.append("return 0;}")
    res = {}
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "probe.c")
        with open(src, "w") as fh:
            fh.write("\n".join(lines))
        for width, flags in (("float", []), ("double", ["-DSLE_USE_DOUBLE"])):
            exe = os.path.join(td, "probe_" + width)
            subprocess.run(["gcc", "-std=c99", *flags, "-I", os.path.dirname(header_path), src, "-o", exe],
                           check=True)
            got = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
            res[width] = {k: int(v) for k, v in (ln.rsplit(" ", 1) for ln in got.splitlines())}
    return res


def probe(header_path: str) -> dict:
    txt = open(header_path).read()
    return {"functions": functions(txt), "structs": structs(txt), "enums": enums(txt),
            "layouts": layouts(header_path)}


def main():
    if not os.path.exists(REFERENCE_HEADER):
        sys.exit("the reference header is not present here; the committed fixture stands")
    data = probe(REFERENCE_HEADER)
    data["source"] = "include/sle_api.h of the reference (v0.20.0)"
    with open(FIXTURE, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    print(f"{len(data['functions'])} functions, {len(data['enums'])} enumerators, "
          f"{len(data['structs'])} structs -> {FIXTURE}")


if __name__ == "__main__":
    main()
