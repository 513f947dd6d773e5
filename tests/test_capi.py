"""The C-ABI library loads and exports every entry point include/sf_hip.h declares (no GPU calls)."""
import ctypes
import os
import re

import sfhip
from sfhip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "sf_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(sf_\w+)\s*\(", text, flags=re.M)))


def test_library_built_in_tree():
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build()"
    assert os.path.dirname(_lib.LIB_PATH).endswith("stokesflow.jl_amd")


def test_exports_every_declared_symbol():
    decl = declared_functions()
    assert len(decl) >= 13
    assert sorted(_lib.EXPORTS) == decl
    raw = ctypes.CDLL(_lib.LIB_PATH)
    for name in decl:
        assert hasattr(raw, name), name


def test_version_and_error_string_without_gpu():
    L = sfhip.load()
    assert L.sf_version() == 100
    # an argument error is reported without touching the device
    h = ctypes.c_void_p()
    assert L.sf_ctx_create(None, 8, 8, None) == _lib.SF_ERR_ARG
    assert b"NULL" in L.sf_last_error()
    assert L.sf_ctx_create(ctypes.byref(h), 2, 8, None) == _lib.SF_ERR_ARG
    assert b">= 3" in L.sf_last_error()
