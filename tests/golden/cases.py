"""Case definitions shared by the golden-fixture generator and the tests.

Every case is a pure function of its arguments (seeded), so the tests rebuild
the exact inputs the upstream program was run on; the fixtures hold outputs
only.
"""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from ixchel2d_amd import case as C  # noqa: E402

GOLDEN_DIR = os.path.dirname(os.path.abspath(__file__))


def nc18x14():
    """Stretched natural convection, small and odd-sized."""
    c = C.natural_convection_case(18, 14, height=1.4, Ra=1e4, iters=20, dt=1.0, seed=1)
    c.params.paq_itera = 10
    c.name = "nc18x14"
    return c


def mixed22x17():
    """Mixed convection with buoyancy in v, multi-segment boundary sides,
    Neumann rows on several sides, non-zero initial u, v, p."""
    mesh = C.make_mesh(22, 17, 1.5, 1.0, "unif")
    prm = C.Params(adimen="mixta", Ra=50.0, Pr=0.71, Ri_1=2.0, dt=0.5, paq_itera=10, itermax=30,
                   rel_pres=0.4, rel_vel=0.6, rel_ener=0.8, flujo_ini="archivo", tempe_ini="archivo")
    S = C.Segment
    big = 1e30
    bcs = {
        "u": C.BoundarySet("u", {"a": [S(-big, big, "diri", 0.0)], "b": [S(-big, big, "diri", 0.0)],
                                 "c": [S(-big, big, "diri", 0.0)],
                                 "d": [S(0.0, 0.7, "diri", 1.0), S(0.7, big, "diri", 0.5)]}),
        "v": C.BoundarySet("v", {"a": [S(-big, big, "diri", 0.0)], "b": [S(-big, big, "diri", 0.0)],
                                 "c": [S(-big, 0.5, "diri", 0.0), S(0.5, big, "neum", 0.0)],
                                 "d": [S(-big, big, "diri", 0.0)]}),
        "t": C.BoundarySet("t", {"a": [S(-big, 0.5, "diri", 1.0), S(0.5, big, "neum", 0.0)],
                                 "b": [S(-big, big, "neum", 0.0)], "c": [S(-big, big, "diri", 0.0)],
                                 "d": [S(0.0, 0.4, "diri", 0.25), S(0.4, 1.1, "neum", 0.1),
                                       S(1.1, big, "diri", 0.5)]}),
    }
    rng = np.random.default_rng(7)
    mi, nj = mesh.mi, mesh.nj
    u0 = np.asfortranarray(0.05 * rng.standard_normal((mi, nj + 1)))
    v0 = np.asfortranarray(0.05 * rng.standard_normal((mi + 1, nj)))
    t0 = np.asfortranarray(0.5 + 0.1 * rng.standard_normal((mi + 1, nj + 1)))
    p0 = np.asfortranarray(0.01 * rng.standard_normal((mi + 1, nj + 1)))
    return C.Case(mesh, prm, bcs, u0, v0, t0, p0, name="mixed22x17")


def ctrl18x14():
    """Finite convergence thresholds and pass limits: exercises upstream's
    inner-loop and SIMPLE-loop exit rules."""
    c = C.natural_convection_case(18, 14, height=1.4, Ra=1e4, iters=10, dt=0.5, seed=2)
    p = c.params
    p.paq_itera, p.itermax = 10, 10
    p.conv_u, p.conv_t, p.conv_p = 0.05, 0.02, 0.12
    p.conv_resi, p.conv_paso = 1e-3, 5e-2
    p.iter_ecuaci_max, p.iter_simple_max = 3, 2
    c.name = "ctrl18x14"
    return c


def lid66():
    """Lid-driven cavity, Re = 100, 65x65 pressure cells, 2000 iterations."""
    c = C.lid_driven_case(66, Re=100.0, iters=2000, dt=1.0)
    c.params.paq_itera = 1000
    c.name = "lid66"
    return c


def nc34():
    """Square natural convection, Ra = 1e5, stretched, for the Nusselt check."""
    c = C.natural_convection_case(34, 34, height=1.0, Ra=1e5, iters=1000, dt=0.01, seed=4)
    c.params.paq_itera = 1000
    c.name = "nc34"
    return c


def nc301x701(iters: int = 5000):
    """The full-size tall cavity: 301x701, tanh beta = 1.2, H/W = 2.33,
    Ra = 1e5, Pr = 0.71, dt = 0.003."""
    c = C.natural_convection_case(301, 701, height=2.33, Ra=1e5, Pr=0.71, beta=1.2, iters=iters,
                                  dt=0.003, seed=0)
    c.params.paq_itera = 1000 if iters % 1000 == 0 else 10
    c.name = "nc301x701"
    return c


def _long_cavity(mi: int, nj: int, width: float, name: str):
    """Mixed convection in a long lid-driven cavity (width x 1, uniform mesh,
    Re = 100, Ri = 0.5, hot wall a, cold wall c, adiabatic b and d), 200
    iterations from rest with T = 0.5 + U(-1e-3, 1e-3)."""
    mesh = C.make_mesh(mi, nj, width, 1.0, "unif")
    prm = C.Params(adimen="mixta", Ra=100.0, Pr=0.71, Ri_1=0.5, dt=0.1, paq_itera=100, itermax=200,
                   rel_pres=0.3, rel_vel=0.7, rel_ener=0.7, flujo_ini="nulo", tempe_ini="archivo")
    bcs = {
        "u": C.uniform_bc("u", ("diri", 0.0), ("diri", 0.0), ("diri", 0.0), ("diri", 1.0)),
        "v": C.uniform_bc("v", ("diri", 0.0), ("diri", 0.0), ("diri", 0.0), ("diri", 0.0)),
        "t": C.uniform_bc("t", ("diri", 1.0), ("neum", 0.0), ("diri", 0.0), ("neum", 0.0)),
    }
    mi, nj = mesh.mi, mesh.nj
    rng = np.random.default_rng(11)
    t0 = np.asfortranarray(0.5 + rng.uniform(-1e-3, 1e-3, size=(mi + 1, nj + 1)))
    z = lambda *sh: np.zeros(sh, order="F")  # noqa: E731
    return C.Case(mesh, prm, bcs, z(mi, nj + 1), z(mi + 1, nj), t0, z(mi + 1, nj + 1), name=name)


# The cases around upstream's own immersed bodies.  Their case files carry no
# body: the reference builds it in definir_cuerpo (each fixture comes from
# the binary built with BODY = its option), and the oracle and the GPU get it
# from with_body().
def ib161x41():
    """161 x 41, 8 x 1, around the square 'cuadr' (side 0.1 centred at
    (6.0, 0.05), on the bottom wall)."""
    return _long_cavity(161, 41, 8.0, "ib161x41")


def ib161x41t():
    """161 x 41, 8 x 1, around the triangle 'trian' (base 0.4 on the bottom
    wall from x = 5.8 to 6.2, apex (6.0, 0.1))."""
    return _long_cavity(161, 41, 8.0, "ib161x41t")


def ib201x41c():
    """201 x 41, 10 x 1, around the rod 'chime' (side 0.3 centred at
    (8.5, 0.5)), which also holds T = 1."""
    return _long_cavity(201, 41, 10.0, "ib201x41c")


# the definir_cuerpo option of the reference build behind a fixture
REF_BODY = {"ib161x41": "cuadr", "ib161x41t": "trian", "ib201x41c": "chime"}


def with_body(case):
    """The case with the immersed body its fixture's reference build defines."""
    if case.name in REF_BODY:
        C.upstream_body_sources(case, REF_BODY[case.name])
    return case


CASES = {"nc18x14": nc18x14, "mixed22x17": mixed22x17, "ctrl18x14": ctrl18x14, "lid66": lid66,
         "nc34": nc34, "nc301x701": nc301x701, "ib161x41": ib161x41, "ib161x41t": ib161x41t,
         "ib201x41c": ib201x41c}

# subsampling of the full-size fixture (every SUB_I-th node in x, SUB_J-th in y)
SUB_I, SUB_J = 6, 10


def fixture_path(name: str) -> str:
    return os.path.join(GOLDEN_DIR, f"{name}.npz")


def load_fixture(name: str):
    with np.load(fixture_path(name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def subsample(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a[::SUB_I, ::SUB_J])


def parse_vtk(data: bytes) -> dict:
    """The sections of a legacy binary VTK structured grid in the layout of
    upstream's postproceso_vtk: the text pieces (bytes) and the big-endian
    FP32 blocks (points, p, t, vel) as arrays."""
    d0 = data.index(b"DIMENSIONS")
    dims = [int(x) for x in data[d0 + 10:data.index(b"\n", d0)].split()]
    n = dims[0] * dims[1] * dims[2]
    j = data.index(b"\n", data.index(b"POINTS ")) + 1
    out = {"header": data[:j], "points": np.frombuffer(data, ">f4", 3 * n, j)}
    j += 12 * n
    for key, marker, cnt in (("p", b"LOOKUP_TABLE default\n", n), ("t", b"LOOKUP_TABLE default\n", n),
                             ("vel", b"VECTORS VELOCITY float\n", 3 * n)):
        k = data.index(marker, j) + len(marker)
        out[key + "_text"] = data[j:k]
        out[key] = np.frombuffer(data, ">f4", cnt, k)
        j = k + 4 * cnt
    if j != len(data):
        raise ValueError("trailing bytes after the VTK velocity block")
    return out
