"""The Fortran host's writers against the files upstream itself wrote: the
VTK sample (postproceso_vtk, after upstream's mod_postproceso.f90:225-300
and the node velocities of main.f90:1295-1313) and the restart files
(postproceso_dat, postproceso_bin).  The final fields of each golden case,
written by the host's writers, give upstream's files byte for byte (VTK:
header, FP32 big-endian points, pressure, temperature, velocities; .dat:
the list-directed header and the D23.15 records).  Runs on the CPU: the
writers are host code and the driver makes no GPU call."""
import os
import subprocess

import numpy as np
import pytest

import cases as K
from ixchel2d_amd import case as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "ixchel2d_amd", "fortran")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
MODS = ("mod_ixchel_hip", "mod_malla", "mod_postproceso", "d_format")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not os.path.exists(FLANG):
        pytest.skip("flang is not installed")
    lib = os.path.join(ROOT, "ixchel2d_amd", "libixchel2d.so")
    assert os.path.exists(lib), "libixchel2d.so not built (run __graft_entry__.build())"
    if not all(os.path.exists(os.path.join(FDIR, m + ".o")) for m in MODS):
        subprocess.run(["make", "-s", "-C", FDIR, f"FC={FLANG}"], check=True)
    out = tmp_path_factory.mktemp("vtk") / "vtk_driver"
    subprocess.run([FLANG, "-O2", "-I", FDIR, "-o", str(out), os.path.join(ROOT, "tests", "vtk_driver.f90")]
                   + [os.path.join(FDIR, m + ".o") for m in MODS]
                   + [f"-L{os.path.dirname(lib)}", "-lixchel2d", "-lm", "-lpthread", f"-Wl,-rpath,{os.path.dirname(lib)}"],
                   check=True, cwd=str(out.parent))
    return str(out)


def write_vtk(driver, name, fields, tmp_path):
    c = K.CASES[name]()
    for f in "uvtp":
        np.ascontiguousarray(fields[f].T, dtype="<f8").tofile(os.path.join(tmp_path, f + ".raw"))
    names = C.write_mesh_files(c, str(tmp_path))
    subprocess.run([driver, *names], check=True, cwd=str(tmp_path), stdout=subprocess.PIPE,
                   stderr=subprocess.STDOUT)
    with open(os.path.join(tmp_path, "out.vtk"), "rb") as fh:
        return fh.read(), c


@pytest.mark.parametrize("name", ["nc18x14", "mixed22x17", "ctrl18x14", "lid66", "nc34"])
def test_vtk_matches_upstream_bytes(driver, name, tmp_path):
    gold = K.load_fixture(name)
    data, _ = write_vtk(driver, name, {f: gold[f] for f in "uvtp"}, tmp_path)
    ref = gold["vtk"].tobytes()
    assert len(data) == len(ref)
    assert data == ref


def write_dat(driver, name, fields, header, tmp_path):
    """out{u,v,p}.dat and out{u,v,p}.bin from the host's writers"""
    c = K.CASES[name]()
    for f in "uvtp":
        np.ascontiguousarray(fields[f].T, dtype="<f8").tofile(os.path.join(tmp_path, f + ".raw"))
    names = C.write_mesh_files(c, str(tmp_path))
    subprocess.run([driver, *names, header], check=True, cwd=str(tmp_path), stdout=subprocess.PIPE,
                   stderr=subprocess.STDOUT)
    out = {}
    for f in "uvp":
        for ext in ("dat", "bin"):
            with open(os.path.join(tmp_path, f"out{f}.{ext}"), "rb") as fh:
                out[f + ext] = fh.read()
    return out, c


@pytest.mark.parametrize("name", ["nc18x14", "mixed22x17", "ctrl18x14", "lid66", "nc34"])
def test_restart_files_match_upstream_bytes(driver, name, tmp_path):
    """postproceso_dat (the D23.15 records formatted in C) and
    postproceso_bin on the reference's final fields give the reference's
    out_*.dat bytes (mod_postproceso.f90 postproceso_dat, main.f90:1329-1354)
    and .bin records (x, y, u / x, y, v / T, p per node, FP64)."""
    gold = K.load_fixture(name)
    head = gold["dat_u"].tobytes().split(b"\n", 1)[0].decode()
    out, c = write_dat(driver, name, {f: gold[f] for f in "uvtp"}, head, tmp_path)
    for f in "uvp":
        assert out[f + "dat"] == gold["dat_" + f].tobytes(), f + ".dat"
    mi, nj = c.mesh.mi, c.mesh.nj
    ub = np.frombuffer(out["ubin"], "<f8").reshape(nj + 1, mi, 3)
    vb = np.frombuffer(out["vbin"], "<f8").reshape(nj, mi + 1, 3)
    pb = np.frombuffer(out["pbin"], "<f8").reshape(nj + 1, mi + 1, 2)
    assert np.array_equal(ub[:, :, 2].T, gold["u"]) and np.array_equal(vb[:, :, 2].T, gold["v"])
    assert np.array_equal(pb[:, :, 0].T, gold["t"]) and np.array_equal(pb[:, :, 1].T, gold["p"])


def test_restart_file_with_a_nonfinite_value(driver, tmp_path):
    """A field holding a NaN goes through Fortran's formatted write instead:
    every other record as before, the NaN record as flang writes it."""
    name = "nc18x14"
    gold = K.load_fixture(name)
    head = gold["dat_u"].tobytes().split(b"\n", 1)[0].decode()
    fields = {f: gold[f].copy() for f in "uvtp"}
    fields["u"][3, 4] = np.nan
    out, c = write_dat(driver, name, fields, head, tmp_path)
    ref = gold["dat_u"].tobytes().split(b"\n")
    got = out["udat"].split(b"\n")
    assert len(got) == len(ref)
    k = 1 + 4 * c.mesh.mi + 3  # header, then records by rows of mi
    bad = [i for i, (a, b) in enumerate(zip(got, ref)) if a != b]
    assert bad == [k]
    assert got[k][:46] == ref[k][:46] and got[k][46:].strip() == b"NaN" and len(got[k]) == 69
    assert out["vdat"] == gold["dat_v"].tobytes() and out["pdat"] == gold["dat_p"].tobytes()


def test_vtk_layout_and_node_velocities(driver, tmp_path):
    """The file parsed back: header keywords and counts, the p-node grid,
    and the node velocities from their definition on random fields
    (interior means of two faces, side columns from the boundary faces,
    bottom and top rows of v from the wall faces, u there 0 apart from the
    corners)."""
    name = "mixed22x17"
    c = K.CASES[name]()
    mi, nj = c.mesh.mi, c.mesh.nj
    rng = np.random.default_rng(3)
    fields = {"u": rng.standard_normal((mi, nj + 1)), "v": rng.standard_normal((mi + 1, nj)),
              "t": rng.standard_normal((mi + 1, nj + 1)), "p": rng.standard_normal((mi + 1, nj + 1))}
    data, c = write_vtk(driver, name, fields, tmp_path)
    npts = (mi + 1) * (nj + 1)
    head = (b"# vtk DataFile Version 2.3\n3D Mesh\nBINARY\nDATASET STRUCTURED_GRID\n"
            + f"DIMENSIONS  {mi + 1} {nj + 1} 1\nPOINTS  {npts} float\n".encode())
    assert data.startswith(head)
    pos = len(head)

    def floats(n):
        nonlocal pos
        a = np.frombuffer(data, dtype=">f4", count=n, offset=pos)
        pos += 4 * n
        return a

    def text(s):
        nonlocal pos
        assert data[pos:pos + len(s)] == s, (data[pos:pos + len(s)], s)
        pos += len(s)

    xp, yp = np.asarray(c.mesh.xp), np.asarray(c.mesh.yp)
    pts = floats(3 * npts).reshape(nj + 1, mi + 1, 3)
    # the mesh files round-trip the coordinates through D23.15 text
    np.testing.assert_allclose(pts[:, :, 0], np.broadcast_to(xp.astype(np.float32), (nj + 1, mi + 1)), rtol=1e-7)
    np.testing.assert_allclose(pts[:, :, 1], np.broadcast_to(yp.astype(np.float32)[:, None], (nj + 1, mi + 1)),
                               rtol=1e-7)
    assert not pts[:, :, 2].any()
    text(f"\nPOINT_DATA  {npts}SCALARS PRESS float\nLOOKUP_TABLE default\n".encode())
    assert np.array_equal(floats(npts), fields["p"].T.ravel().astype(np.float32))
    text(b"\nSCALARS TEMPER float\nLOOKUP_TABLE default\n")
    assert np.array_equal(floats(npts), fields["t"].T.ravel().astype(np.float32))
    text(b"\nVECTORS VELOCITY float\n")
    vel = floats(3 * npts).reshape(nj + 1, mi + 1, 3)
    assert pos == len(data)
    u, v = fields["u"], fields["v"]
    uf = np.zeros((mi + 1, nj + 1))
    vf = np.zeros((mi + 1, nj + 1))
    for j in range(1, nj):              # 0-based interior rows 1..nj-1
        for i in range(1, mi):
            uf[i, j] = (u[i, j] + u[i - 1, j]) / 2
            vf[i, j] = (v[i, j] + v[i, j - 1]) / 2
        vf[0, j] = v[0, j]
        vf[mi, j] = v[mi, j]
    uf[0, :] = u[0, :]
    uf[mi, :] = u[mi - 1, :]
    vf[:, 0] = v[:, 0]
    vf[:, nj] = v[:, nj - 1]
    assert np.array_equal(vel[:, :, 0], uf.T.astype(np.float32))
    assert np.array_equal(vel[:, :, 1], vf.T.astype(np.float32))
    assert not vel[:, :, 2].any()
    assert not uf[1:mi, 0].any() and not uf[1:mi, nj].any()
