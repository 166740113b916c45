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


@pytest.mark.parametrize("mi,nj", [(3, 10), (10, 3)])
def test_meshes_below_the_smallest_are_rejected(mi, nj):
    """ixh_create refuses mi or nj below 4 with its error code and message
    before it touches the GPU (the grid is handed over directly: the Python
    mesh builder refuses such meshes itself)."""
    import numpy as np
    from ixchel2d_amd import case as C
    from ixchel2d_amd.simple import params_of
    case = C.natural_convection_case(10, 10, height=1.0, Ra=1e4, iters=1, dt=0.05, seed=5)
    m = case.mesh
    geo = [np.ascontiguousarray(x, dtype=np.float64) for x in (m.xu, m.yv, m.xp, m.yp)]
    grid = _lib.Grid(mi, nj, *[x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for x in geo])
    prm = params_of(case, "pcr", "fp64")
    h = ctypes.c_void_p()
    L = _lib.load()
    assert L.ixh_create(ctypes.byref(grid), ctypes.byref(prm), ctypes.byref(h)) == -1  # IXH_ERR_ARG
    assert not h.value
    assert "mesh needs mi, nj >= 4" in L.ixh_last_error().decode()
