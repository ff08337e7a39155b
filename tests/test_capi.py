"""The C-ABI library loads without a GPU and exports every entry point that
include/mpm.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "mpm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[a-z0-9_]+)\s*\(", text)))


def _lib():
    from code_mpm_amd import capi
    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libmpm_amd.so not built")
    return ctypes.CDLL(capi.LIB_PATH)


def test_every_declared_symbol_is_exported():
    lib = _lib()
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from code_mpm_amd import capi
    assert sorted(capi.SIGNATURES) == _declared()


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from code_mpm_amd import capi
    lib = capi.load()
    h = ctypes.c_void_p()
    rc = lib.mpm_create(ctypes.byref(h), 0)
    assert rc != 0 and not h.value
    assert lib.mpm_last_error(None)


def test_null_engine_is_an_argument_error():
    from code_mpm_amd import capi
    lib = capi.load()
    assert lib.mpm_step(None, 1.0, 1) == -1
    assert lib.mpm_destroy(None) == 0
