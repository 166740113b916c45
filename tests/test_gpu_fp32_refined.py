"""FP32-refined path (precision IXH_FP32_REFINED): every field stays FP64;
each line sweep assembles its rows in FP32 and solves, in FP32, for the
increment of the line against the line's FP64 residual.  That residual takes
a_P only through neighbour differences, the sources at the FP64 cell area,
and the time term as phi_ant - phi, so the SIMPLE iteration refines the FP64
fields: it converges to the steady state of the FP32-rounded neighbour
coefficients, within about 1e-8 of the FP64 one, instead of stalling where
the FP32 update falls below the FP32 spacing of the fields.  Bounds are
about 5x the errors measured on MI355X:
  - upstream fixtures nc18x14, mixed22x17, lid66, nc34: fields 1.2e-8 ..
    2.4e-7 (FP32-mixed: 3e-7 .. 4e-5);
  - a 34x34 natural-convection cavity run to the steady state (dt = 1,
    12000 iterations): residual 2.9e-10, hot-wall Nu 1.4e-9 from FP64,
    fields 2.9e-8 (FP32-mixed: residual 4.4e-9, Nu 4.8e-6, fields 4.6e-6)."""
import numpy as np
import pytest

import cases as K
from ixchel2d_amd import case as C

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def _cavity(case, precision="fp32_refined"):
    from ixchel2d_amd.simple import Cavity
    return Cavity(case, solver="pcr", precision=precision)


TOL = {"nc18x14": 3e-7, "mixed22x17": 1.2e-6, "lid66": 1e-7, "nc34": 5e-7}


@pytest.mark.parametrize("name", list(TOL))
def test_fp32_refined_against_upstream_fixture(name):
    case = K.CASES[name]()
    gold = K.load_fixture(name)
    g = _cavity(case)
    h = g.run(case.params.itermax)
    for f in "uvtp":
        a = g.get(f)
        assert np.all(np.isfinite(a))
        if np.max(np.abs(gold[f])) == 0.0:
            assert np.max(np.abs(a)) == 0.0
            continue
        assert relerr(a, gold[f]) <= TOL[name], f"{name}: {f} rel err {relerr(a, gold[f]):.2e}"
    np.testing.assert_allclose(h["maxbo"][:10], gold["history"][:10, 0], rtol=5e-7)


def test_fp32_refined_reaches_the_fp64_steady_state():
    case = C.natural_convection_case(34, 34, height=1.0, Ra=1e5, iters=100, dt=1.0, seed=4)
    g64 = _cavity(case, "fp64")
    g = _cavity(case)
    g64.run(12000)
    h = g.run(12000)
    assert h["residuo"][-1] <= 1.5e-9
    nu, nu64 = g.nusselt_x()[0], g64.nusselt_x()[0]
    assert abs(nu - nu64) / abs(nu64) <= 1e-7
    for f in "uvtp":
        assert relerr(g.get(f), g64.get(f)) <= 1.5e-7, f


def test_fp32_refined_immersed_cylinder():
    case = C.natural_convection_case(40, 48, height=1.2, Ra=1e5, iters=30, dt=0.01, seed=5)
    C.cylinder_sources(case, 0.55, 0.6, 0.15, t_body=1.0)
    g = _cavity(case)
    g.run(30)
    g64 = _cavity(case, "fp64")
    g64.run(30)
    m = case.mesh
    X, Y = np.meshgrid(m.xp, m.yp, indexing="ij")
    inside = (X - 0.55) ** 2 + (Y - 0.6) ** 2 <= 0.15 ** 2
    for f in "uvtp":
        assert np.all(np.isfinite(g.get(f))), f
    # a penalised row's increment is alpha (Sc/-Sp - phi), the step of
    # upstream's under-relaxed penalty row: T inside follows the FP64 path's
    assert np.max(np.abs(g.get("t")[inside] - g64.get("t")[inside])) <= 1e-12
    for f in "uvt":
        assert relerr(g.get(f), g64.get(f)) <= 1.5e-7, f
    p32, p64 = g.get("p"), g64.get("p")
    assert relerr(p32[~inside], p64[~inside]) <= 1.5e-7


def test_fp32_refined_batch_matches_single_cavity():
    from ixchel2d_amd.simple import Batch
    cases = [C.lid_driven_case(30, Re=400.0, iters=10), K.nc34(),
             C.natural_convection_case(26, 40, height=1.5, Ra=1e5, iters=10, dt=0.01, seed=4)]
    C.cylinder_sources(cases[0], 0.5, 0.5, 0.12, t_body=1.0)
    cav = [_cavity(c) for c in cases]
    b = Batch(cav, order="xcd", graph=True)
    b.run(12)
    for c, case in enumerate(cases):
        alone = _cavity(case)
        alone.run(12)
        for f in "uvtp":
            assert np.array_equal(cav[c].get(f), alone.get(f)), f"cavity {c}: {f}"
    with pytest.raises(Exception):
        Batch([cav[0], _cavity(cases[1], "fp32_mixed")])


def test_fp32_refined_rejects_other_solvers_and_the_adjoint():
    from ixchel2d_amd import _lib
    from ixchel2d_amd.simple import Cavity
    with pytest.raises(_lib.IxhError):
        Cavity(K.nc18x14(), solver="thomas", precision="fp32_refined")
    g = _cavity(K.nc18x14())
    with pytest.raises(_lib.IxhError):
        g.adjoint_record(20)
