"""The C-ABI library loads and exports every entry point include/fdtd_hip.h
declares (no device work: only calls that fail before touching HIP)."""

import ctypes
import os
import re

from fdtd_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "fdtd_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fdtd_\w+)\s*\(", text)))


def test_library_is_built_in_tree():
    assert os.path.exists(_native.LIB_PATH), "run __graft_entry__.build() first"


def test_every_declared_symbol_is_exported_and_bound():
    lib = _native.lib()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert names == _native.exported_symbols()


def test_single_hip_runtime_mapped():
    _native.lib()
    assert len(_native._hip_runtimes_mapped()) == 1


def test_version_and_argument_errors():
    lib = _native.lib()
    assert b"gfx950" in lib.fdtd_version()
    out = ctypes.c_void_p()
    rc = lib.fdtd_plan_create(0, 4, 4, 4, 0.5, None, None, None, None, ctypes.byref(out))
    assert rc != 0 and b"null" in lib.fdtd_last_error()
    rc = lib.fdtd_plan_create(7, 4, 4, 4, 0.5, None, None, None, None, ctypes.byref(out))
    assert rc != 0 and b"dtype" in lib.fdtd_last_error()
    assert lib.fdtd_half_step_E(None, 0, None) != 0
    assert lib.fdtd_plan_destroy(None) == 0
