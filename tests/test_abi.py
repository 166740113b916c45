"""The C-ABI library loads and exports every symbol include/ixchel2d.h
declares; the bindings name the same set.  No GPU call is made."""
import ctypes
import os
import re

import pytest

from ixchel2d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "ixchel2d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ixh_[a-z_]+)\s*\(", text)))


def test_header_declares_the_bound_symbols():
    assert header_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    missing = [s for s in header_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert L.ixh_version().startswith(b"ixchel2d-amd")


def test_fortran_binding_names_every_symbol():
    f = open(os.path.join(ROOT, "ixchel2d_amd", "fortran", "mod_ixchel_hip.f90")).read()
    bound = set(re.findall(r"bind\(C,\s*name\s*=\s*['\"](ixh_[a-z_]+)['\"]\)", f))
    assert bound == set(header_symbols())


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.IxhError):
        _lib.load(str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    _lib.load()


def test_errors_come_back_as_codes_without_a_gpu():
    L = _lib.load()
    h = ctypes.c_void_p()
    # NULL grid: rejected before any device work
    assert L.ixh_create(None, None, ctypes.byref(h)) == -1
    assert b"NULL" in L.ixh_last_error()
