"""The Fortran host (ixchel2d_amd/fortran/ixchel2d_gpu) as a drop-in for the
upstream program: the same case directory in, the same fort.102 history,
the same exact .bin field dumps and the same VTK sample out (Thomas line
solver: bit-identical; PCR: 1e-10 relative)."""
import os
import subprocess

import numpy as np
import pytest

import cases as K
from ixchel2d_amd import case as C

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ixchel2d_amd", "fortran", "ixchel2d_gpu")


@pytest.mark.parametrize("name", ["nc18x14", "mixed22x17", "ctrl18x14", "nc34"])
@pytest.mark.parametrize("solver", ["thomas", "pcr"])
def test_fortran_host_reproduces_upstream_outputs(name, solver, tmp_path):
    assert os.path.exists(EXE), "Fortran host not built"
    c = K.CASES[name]()
    gold = K.load_fixture(name)
    C.write_reference_case(c, str(tmp_path))
    env = dict(os.environ, IXCHEL2D_SOLVER=solver)
    r = subprocess.run([EXE], cwd=tmp_path, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    out = C.read_reference_outputs(c, str(tmp_path))
    for f in "uvtp":
        if solver == "thomas":
            assert np.array_equal(out[f], gold[f]), f
        else:
            rel = np.max(np.abs(out[f] - gold[f])) / max(np.max(np.abs(gold[f])), 1e-300)
            assert rel <= 1e-10, (f, rel)
    assert out["history"].shape == gold["history"].shape
    np.testing.assert_allclose(out["history"], gold["history"], rtol=1e-9, atol=1e-15)
    if gold["nusselt"].size:
        np.testing.assert_allclose(out["nusselt"][:, :3], gold["nusselt"][:, :3], rtol=1e-12, atol=1e-16)
    # the last VTK sample: upstream's bytes (Thomas); with PCR the FP32
    # values within one FP32 rounding of upstream's
    vdir = os.path.join(tmp_path, f"n{c.mesh.nj:05d}m{c.mesh.mi:05d}R{c.params.rec()}")
    with open(os.path.join(vdir, str(gold["vtk_name"])), "rb") as fh:
        data = fh.read()
    ref = gold["vtk"].tobytes()
    if solver == "thomas":
        assert data == ref
    else:
        a, b = K.parse_vtk(data), K.parse_vtk(ref)
        assert a.keys() == b.keys()
        for k in a:
            if k in ("points", "p", "t", "vel"):
                np.testing.assert_allclose(a[k], b[k], rtol=2.5e-7, atol=1e-9 * float(np.max(np.abs(b[k]))), err_msg=k)
            else:
                assert a[k] == b[k], k


@pytest.mark.parametrize("precision,bound", [("fp32_mixed", 2e-4), ("fp32_refined", 5e-7)])
def test_fortran_host_fp32_paths(precision, bound, tmp_path):
    """IXCHEL2D_PRECISION selects an FP32 path: the same outputs within that
    path's bound on the nc34 fixture after 1000 iterations (FP32-mixed 2e-4,
    FP32-refined 5e-7 relative)."""
    assert os.path.exists(EXE), "Fortran host not built"
    c = K.nc34()
    gold = K.load_fixture("nc34")
    C.write_reference_case(c, str(tmp_path))
    env = dict(os.environ, IXCHEL2D_SOLVER="pcr", IXCHEL2D_PRECISION=precision)
    r = subprocess.run([EXE], cwd=tmp_path, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    out = C.read_reference_outputs(c, str(tmp_path))
    for f in "uvtp":
        rel = np.max(np.abs(out[f] - gold[f])) / max(np.max(np.abs(gold[f])), 1e-300)
        assert rel <= bound, (f, rel)
