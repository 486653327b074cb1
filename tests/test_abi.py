"""The C-ABI library loads without a GPU and exports every function that
include/fwi.h declares (no compute calls here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fwi.h")
LIB = os.path.join(ROOT, "fwi_amd", "libfwi_hip.so")


def declared_functions():
    """Names of the functions declared in include/fwi.h (one declaration per `;`)."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"^\s*#.*$", "", text, flags=re.M)
    text = re.sub(r"typedef\s+struct\s+\w+\s*\{.*?\}\s*\w+\s*;", "", text, flags=re.S)
    text = re.sub(r"enum\s*\{.*?\}\s*;", "", text, flags=re.S)
    names = set()
    for stmt in text.split(";"):
        m = re.search(r"([A-Za-z_][A-Za-z0-9_]*)\s*\(", stmt)
        if m and "typedef" not in stmt:
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_dropin_symbols():
    names = declared_functions()
    for n in ("solver_acoustic_ss_", "solver_forward_simple_", "fwi_create", "fwi_gradient", "fwi_destroy"):
        assert n in names


def test_library_loads_and_exports_all_declared_symbols():
    if not os.path.exists(LIB):
        pytest.fail("fwi_amd/libfwi_hip.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(declared_functions()) <= exported


def test_library_has_gfx950_code_object():
    """The library (built here) carries a HIP fat binary with a gfx950 code object."""
    out = subprocess.run(["readelf", "-S", "-W", LIB], capture_output=True, text=True).stdout
    assert ".hip_fatbin" in out
    assert b"amdgcn-amd-amdhsa--gfx950" in open(LIB, "rb").read()


def test_last_error_without_gpu_work():
    lib = C.CDLL(LIB)
    lib.fwi_last_error.restype = C.c_char_p
    assert isinstance(lib.fwi_last_error(), bytes)
