"""The C-ABI library loads and exports every entry point include/gdx.h declares (no GPU
compute here), and argument validation fails with a status + message before any launch."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "gdx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gdx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from graddft_amd import _lib
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the binding table covers the same set
    assert set(_lib.SIGNATURES) == set(names)


def test_validation_errors_are_reported_without_a_gpu():
    from graddft_amd import _lib
    lib = _lib.load()
    d = _lib.XCDesc()
    d.npts, d.npts_pad, d.nao, d.nao_pad = 100, 100, 10, 16   # npts_pad not a multiple of 64
    d.kind, d.precision = _lib.GDX_FUNC_B3LYP, _lib.GDX_FP64
    nbytes = ctypes.c_size_t(0)
    rc = lib.gdx_xc_workspace_size(ctypes.byref(d), ctypes.byref(nbytes))
    assert rc == -1
    assert b"multiple of 64" in lib.gdx_last_error()
    d.npts_pad = 128
    d.precision = _lib.GDX_BF16X6 + 5
    assert lib.gdx_xc_workspace_size(ctypes.byref(d), ctypes.byref(nbytes)) == -3
    d.precision = _lib.GDX_FP64
    assert lib.gdx_xc_workspace_size(ctypes.byref(d), ctypes.byref(nbytes)) == 0
    assert nbytes.value > 0


def test_ints_library_loads():
    from graddft_amd.interface import integrals
    lib = integrals._lib()
    assert hasattr(lib, "gdx_int1e") and hasattr(lib, "gdx_int2e_packed")
