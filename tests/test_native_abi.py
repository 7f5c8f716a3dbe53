"""The C-ABI library loads and exports every entry point declared in
``include/cmhip.h`` (CPU only: no compute calls)."""

import os
import re

import pytest

from contactmechanics_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "cmhip.h")) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared_functions()
    for required in ["cm_ctx_create", "cm_plan_create", "cm_evaluate_disp", "cm_evaluate_force",
                     "cm_ccg_solve", "cm_ccg_solve_batch", "cm_last_error"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("libcmhip.so not built: run __graft_entry__.build()")
    lib = _native.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the whole header
    assert set(declared_functions()) <= set(_native.SIGNATURES)
    assert lib.cm_version().decode().startswith("cmhip")


def test_ctypes_struct_layouts_match_header(tmp_path):
    """Sizes and field offsets of the ctypes structs equal the C compiler's
    for include/cmhip.h."""
    import ctypes
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    structs = {"cm_substrate_desc": _native.SubstrateDesc, "cm_ccg_params": _native.CcgParams,
               "cm_ccg_stats": _native.CcgStats}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "cmhip.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(line.split()[:2]): int(line.split()[2]) for line in out if line.strip()}
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
