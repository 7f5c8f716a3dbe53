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


def test_ctypes_struct_layouts_match_header():
    import ctypes
    # int,int,int,int,double,double,double,double,int,(pad)double,double
    assert ctypes.sizeof(_native.SubstrateDesc) == 4 * 4 + 4 * 8 + 8 + 2 * 8
    assert ctypes.sizeof(_native.CcgStats) == 4 * 4 + 14 * 8
