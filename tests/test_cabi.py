"""The C-ABI library loads and exports every symbol include/ahm.h declares
(no GPU calls)."""
import os
import re

from conftest import ROOT


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ahm.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ahm_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    from ahmed_amd import _lib
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(lib, s), s
        assert s in _lib.SIGNATURES, s


def test_last_error_without_gpu():
    from ahmed_amd import _lib
    lib = _lib.load()
    assert isinstance(lib.ahm_last_error(), bytes)
    # a NULL handle is rejected before any device work
    rc = lib.ahm_hmat_stats_get(None, None)
    assert rc == 1
    assert b"NULL" in lib.ahm_last_error()
