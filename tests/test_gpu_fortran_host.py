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
    # the restart files of the last package (main.f90:1329-1354): upstream's
    # bytes with Thomas, the same numbers to 1e-10 with PCR
    base = os.path.join(tmp_path, f"out_n{c.mesh.nj:05d}m{c.mesh.mi:05d}_R{c.params.rec()}")
    for f in "uvp":
        with open(base + f + ".dat", "rb") as fh:
            data = fh.read()
        ref = gold["dat_" + f].tobytes()
        if solver == "thomas":
            assert data == ref, f + ".dat"
        else:
            a, b = _dat_numbers(data), _dat_numbers(ref)
            assert a.shape == b.shape
            np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-10 * float(np.max(np.abs(b))), err_msg=f)
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


def _dat_numbers(data: bytes) -> np.ndarray:
    """The values of an out_*.dat file (header line, then D23.15 records)."""
    lines = data.decode().splitlines()
    vals = [float(x) for ln in lines[1:] for x in ln.replace("D", "E").split()]
    return np.array(vals)


@pytest.mark.parametrize("precision,bound", [("fp32_mixed", 2e-4), ("fp32_refined", 5e-7),
                                             ("fp32_stochastic", 2e-6)])
def test_fortran_host_fp32_paths(precision, bound, tmp_path):
    """IXCHEL2D_PRECISION selects an FP32 path: the same outputs within that
    path's bound on the nc34 fixture after 1000 iterations (FP32-mixed 2e-4,
    FP32-refined 5e-7, FP32-stochastic 2e-6 relative; the outputs pass
    through the D23.15 text of the restart files)."""
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


@pytest.mark.parametrize("name", ["nc18x14", "mixed22x17", "ctrl18x14", "nc34"])
def test_fortran_phase_procedures_reproduce_upstream(name, tmp_path):
    """IXCHEL2D_LOOP=phases: every time step runs through ecuacion_velu,
    ecuacion_velv, inicializa_corrector_presion, ecuacion_corr_pres,
    corrige_presion_velocidad, ecuacion_energia and residuo_u with upstream's
    loops and exit rules written in Fortran (main.f90:395-1270); with the
    Thomas solver the fields and fort.102 equal upstream's (ctrl18x14 takes
    the multi-pass inner loops and the SIMPLE exits)."""
    assert os.path.exists(EXE), "Fortran host not built"
    c = K.CASES[name]()
    gold = K.load_fixture(name)
    C.write_reference_case(c, str(tmp_path))
    env = dict(os.environ, IXCHEL2D_SOLVER="thomas", IXCHEL2D_LOOP="phases")
    r = subprocess.run([EXE], cwd=tmp_path, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    out = C.read_reference_outputs(c, str(tmp_path))
    for f in "uvtp":
        assert np.array_equal(out[f], gold[f]), f
    # the norms are folded from per-line partials in a fixed order, not
    # summed serially as upstream does: equal to the last bits
    assert out["history"].shape == gold["history"].shape
    np.testing.assert_allclose(out["history"], gold["history"], rtol=1e-12, atol=1e-300)
    if gold["nusselt"].size:
        np.testing.assert_allclose(out["nusselt"][:, :3], gold["nusselt"][:, :3], rtol=1e-12, atol=1e-16)


def test_fortran_tridiagonal_and_tri(tmp_path):
    """tridiagonal (mod_solucionador.f90:25-47) and tri (tri.f90:8-25, the
    sub-diagonal stored one place lower) through the Fortran module, one
    line per call: each system equals the oracle's Thomas solve bit for bit
    and solves the dense system."""
    from oracle.pyoracle import tridiagonal as otri
    drv = os.path.join(ROOT, "tests", "tri_driver")
    assert os.path.exists(drv), "tests/tri_driver not built"
    rng = np.random.default_rng(5)
    for n in (2, 5, 64, 301, 702):
        nsys = 4
        a = rng.uniform(-1.0, 0.0, (nsys, n))
        c = rng.uniform(-1.0, 0.0, (nsys, n))
        b = 2.5 + rng.uniform(0.0, 1.0, (nsys, n))
        rr = rng.standard_normal((nsys, n))
        with open(tmp_path / "sistemas.raw", "wb") as fh:
            fh.write(np.array([n, nsys], dtype=np.int32).tobytes())
            for arr in (a, b, c, rr):
                fh.write(np.ascontiguousarray(arr).tobytes())  # row k = system k = Fortran column k
        run = subprocess.run([drv], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                             timeout=60)
        assert run.returncode == 0, run.stdout[-2000:]
        res = np.fromfile(tmp_path / "soluciones.raw", dtype=np.float64).reshape(4, nsys, n)
        for k in range(nsys):
            xo, bo = otri(a[k], b[k], c[k], rr[k])
            assert np.array_equal(res[0, k], xo), (n, k, "tridiagonal")
            assert np.array_equal(res[1, k], bo), (n, k, "tridiagonal b")
            assert np.array_equal(res[2, k], xo), (n, k, "tri")
            assert np.array_equal(res[3, k], bo), (n, k, "tri b")
            A = np.diag(b[k]) + np.diag(a[k, 1:], -1) + np.diag(c[k, :-1], 1)
            np.testing.assert_allclose(A @ res[0, k], rr[k], rtol=1e-12, atol=1e-12)


def _run_host(case, workdir, **env):
    e = dict(os.environ, IXCHEL2D_SOLVER="thomas", **env)
    r = subprocess.run([EXE], cwd=workdir, env=e, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    return C.read_reference_outputs(case, str(workdir))


def test_fortran_host_restart_round_trip(tmp_path):
    """Checkpoint and resume through the restart files: 4 packages of 10
    iterations uninterrupted, against 2 packages, then a second run that reads
    the first run's out_*.dat files as its mesh and initial fields
    (flujo_ini = tempe_ini = 'archivo', main.f90:1329-1354 and
    mod_malla.f90:137-162) for the last 2.  The header carries the iteration
    count (20), which the resumed run continues from.  The files hold D23.15
    text, 16 significant digits, so the resumed fields agree to 1e-12
    relative, and the resumed fort.102 to 1e-9."""
    import shutil
    c = K.nc18x14()
    c.params.paq_itera = 10
    c.params.itermax = 40
    whole = tmp_path / "whole"
    C.write_reference_case(c, str(whole))
    ref = _run_host(c, whole)
    first = tmp_path / "first"
    c.params.itermax = 20
    C.write_reference_case(c, str(first))
    _run_host(c, first)
    second = tmp_path / "second"
    second.mkdir()
    tag = f"out_n{c.mesh.nj:05d}m{c.mesh.mi:05d}_R{c.params.rec()}"
    names = []
    for f in "uvp":
        shutil.copy(first / (tag + f + ".dat"), second / f"restart_{f}.dat")
        names.append(f"restart_{f}.dat")
    with open(first / (tag + "u.dat")) as fh:
        header = fh.readline().split()
    assert int(header[2]) == 20
    c.params.flujo_ini = "archivo"
    c.params.tempe_ini = "archivo"
    C.write_reference_case(c, str(second))
    lines = open(second / "parametros.dat").read().splitlines()
    lines[17:20] = names
    (second / "parametros.dat").write_text("\n".join(lines) + "\n")
    out = _run_host(c, second)
    for f in "uvtp":
        rel = np.max(np.abs(out[f] - ref[f])) / max(np.max(np.abs(ref[f])), 1e-300)
        assert rel <= 1e-12, (f, rel)
    np.testing.assert_allclose(out["history"], ref["history"][20:], rtol=1e-9, atol=1e-15)
    # the resumed run's last restart files carry the total count
    with open(second / (tag + "u.dat")) as fh:
        assert int(fh.readline().split()[2]) == 40


@pytest.mark.parametrize("name", ["ib161x41", "ib161x41t", "ib201x41c"])
def test_fortran_host_immersed_bodies_against_upstream(tmp_path, name):
    """IXCHEL2D_CUERPO=cuadr | trian | chime: definir_cuerpo of the Fortran
    host (ixchel2d_amd/fortran/mod_front_inmersa.f90) against the reference
    built with the same option (mod_front_inmersa.f90:54-179, main.f90:327):
    with the Thomas solver the fields, the restart files and the VTK sample
    equal upstream's bit for bit."""
    assert os.path.exists(EXE), "Fortran host not built"
    c = K.CASES[name]()
    gold = K.load_fixture(name)
    C.write_reference_case(c, str(tmp_path))
    out = _run_host(c, tmp_path, IXCHEL2D_CUERPO=K.REF_BODY[name])
    for f in "uvtp":
        assert np.array_equal(out[f], gold[f]), f
    np.testing.assert_allclose(out["history"], gold["history"], rtol=1e-12, atol=1e-300)
    base = os.path.join(tmp_path, f"out_n{c.mesh.nj:05d}m{c.mesh.mi:05d}_R{c.params.rec()}")
    for f in "uvp":
        with open(base + f + ".dat", "rb") as fh:
            assert fh.read() == gold["dat_" + f].tobytes(), f + ".dat"
    vdir = os.path.join(tmp_path, f"n{c.mesh.nj:05d}m{c.mesh.mi:05d}R{c.params.rec()}")
    with open(os.path.join(vdir, str(gold["vtk_name"])), "rb") as fh:
        assert fh.read() == gold["vtk"].tobytes()
