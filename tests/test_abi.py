"""The C-ABI library without a GPU: it loads, exports every symbol that
include/bmg.h declares, and refuses to run (instead of falling back to the
CPU) when no device is present."""
import ctypes
import os
import re
import subprocess

import pytest

import basilisk_amd as bmg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "bmg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bmg_[a-z_]+)\s*\(", text)))


def test_library_loads_and_exports_header():
    lib = bmg.library()
    declared = header_symbols()
    assert declared, "no declarations found"
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(bmg.EXPORTS) == declared


def test_dynamic_symbol_table_matches_header():
    out = subprocess.run(["nm", "-D", "--defined-only", bmg.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = sorted(l.split()[-1] for l in out.splitlines() if l.split()[-1].startswith("bmg_"))
    assert exported == header_symbols()


def test_struct_layouts():
    # mgstats, poisson.h:108-114: int, 3 doubles, 2 ints
    assert ctypes.sizeof(bmg.MGStats) == 40
    assert bmg.MGStats.resb.offset == 8 and bmg.MGStats.nrelax.offset == 32


def test_version_and_errors_without_device():
    lib = bmg.library()
    assert b"gfx950" in lib.bmg_version()
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    g = ctypes.c_void_p()
    rc = lib.bmg_grid_new(ctypes.byref(g), 3, 4, 1.0, 0.0, 0.0, 0.0)
    assert rc == 2  # BMG_ERR_HIP: no silent CPU fallback
    assert b"GPU" in lib.bmg_last_error() or b"HIP" in lib.bmg_last_error()
    with pytest.raises(bmg.BMGError):
        bmg.Grid(2, 4)


def test_argument_validation_without_device():
    lib = bmg.library()
    g = ctypes.c_void_p()
    assert lib.bmg_grid_new(ctypes.byref(g), 4, 4, 1.0, 0.0, 0.0, 0.0) == 1
    assert lib.bmg_grid_new(ctypes.byref(g), 3, 11, 1.0, 0.0, 0.0, 0.0) == 1
    assert lib.bmg_grid_new(None, 3, 4, 1.0, 0.0, 0.0, 0.0) == 1
    assert lib.bmg_set_precision(None, 0) == 1
