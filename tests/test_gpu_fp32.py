"""FP32-mixed path (precision IXH_FP32_MIXED): sweeps and coefficient
assembly in FP32 with FP32 storage of u, v, T, p', a_u, a_v; mass residual,
p' right-hand side, p, the norms, the u residual and the Nusselt integrals in
FP64.  Its error against the FP64 upstream fixtures is bounded by the FP32
rounding of the fields, accumulated over the iterations; the bounds below are
a few times the errors measured on MI355X:
  - nc18x14, mixed22x17 (20-30 iterations): fields 1e-6 .. 2e-6;
  - lid66 (2000 iterations): 4e-6;
  - nc34 (1000 iterations, transient): 4e-5;
  - nc301x701 (5000 iterations, transient): fields 1.6e-4 (bound 4e-4),
    hot-wall Nu 4.5e-4 (bound 1e-3) -- above the 1e-4 Nu target, which the
    FP32-refined and FP32-stochastic paths meet (their own test files).
Immersed bodies become rows phi = alpha Sc/-Sp + (1 - alpha) phi_old, the
limit of upstream's under-relaxed 1e51 penalty row (which does not fit
FP32): the body follows the FP64 path's values."""
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


def _cavity(case, precision="fp32_mixed"):
    from ixchel2d_amd.simple import Cavity
    return Cavity(case, solver="pcr", precision=precision)


TOL = {"nc18x14": 5e-6, "mixed22x17": 1e-5, "lid66": 2e-5, "nc34": 2e-4}


@pytest.mark.parametrize("name", list(TOL))
def test_fp32_mixed_against_upstream_fixture(name):
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
    # FP64 norms of FP32 fields: close to upstream's while far above the FP32 floor
    np.testing.assert_allclose(h["maxbo"][:10], gold["history"][:10, 0], rtol=1e-4)


def test_fp32_mixed_full_size_5000_iterations():
    case = K.nc301x701()
    gold = K.load_fixture("nc301x701")
    g = _cavity(case)
    g.run(5000)
    # measured on MI355X: fields within 1.2e-4 .. 1.6e-4 of their maxima,
    # hot-wall Nu 4.5e-4 from FP64.  This path does not meet the 1e-4 Nu
    # bound (FP32 storage of fields that change by less than their FP32
    # spacing per step, rounded to nearest); the FP32-refined and
    # FP32-stochastic paths do and are tested to it
    for f in "uvtp":
        a = g.get(f)
        assert relerr(K.subsample(a), gold[f + "_sub"]) <= 4e-4, f
        assert relerr(a[a.shape[0] // 2, :], gold[f + "_xmid"]) <= 4e-4, f
    g64 = _cavity(case, "fp64")
    g64.run(5000)
    nu32, nu64 = g.nusselt_x()[0], g64.nusselt_x()[0]
    assert abs(nu32 - nu64) / abs(nu64) <= 1e-3


def test_fp32_mixed_nusselt_is_fp64_integral_of_fp32_field():
    """The Nusselt kernels read the FP32 temperature and integrate in FP64:
    the result equals the oracle's integral of the same (rounded) field."""
    from oracle.pyoracle import Oracle
    case = K.nc34()
    g = _cavity(case)
    g.run(200)
    o = Oracle(case)
    for f in ("u", "v", "t", "p"):
        o.set(f, g.get(f))
    assert g.nusselt_y() == o.nusselt_y()
    assert g.nusselt_x() == o.nusselt_x()


def test_fp32_mixed_immersed_cylinder():
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
    # a penalised row relaxes toward the body value as upstream's
    # under-relaxed penalty row does: T inside follows the FP64 path's
    assert np.max(np.abs(g.get("t")[inside] - g64.get("t")[inside])) <= 1e-6
    for f in "uvt":
        assert relerr(g.get(f), g64.get(f)) <= 2e-5, f
    p32, p64 = g.get("p"), g64.get("p")
    assert relerr(p32[~inside], p64[~inside]) <= 2e-5


def test_fp32_mixed_batch_matches_single_cavity():
    from ixchel2d_amd.simple import Batch
    cases = [C.lid_driven_case(30, Re=400.0, iters=10), K.nc34(),
             C.natural_convection_case(26, 40, height=1.5, Ra=1e5, iters=10, dt=0.01, seed=4)]
    C.cylinder_sources(cases[0], 0.5, 0.5, 0.12, t_body=1.0)
    cav = [_cavity(c) for c in cases]
    b = Batch(cav, order="xcd")
    b.run(12)
    for c, case in enumerate(cases):
        alone = _cavity(case)
        alone.run(12)
        for f in "uvtp":
            assert np.array_equal(cav[c].get(f), alone.get(f)), f"cavity {c}: {f}"
    with pytest.raises(Exception):
        Batch([cav[0], _cavity(cases[1], "fp64")])


def test_fp32_mixed_rejects_other_solvers():
    from ixchel2d_amd import _lib
    from ixchel2d_amd.simple import Cavity
    with pytest.raises(_lib.IxhError):
        Cavity(K.nc18x14(), solver="thomas", precision="fp32_mixed")


# cavities of the 256-cavity batch workload (seed 0) whose FP32 line solves
# once went infinite (first at the step given): lid-driven at Re 2200-3100,
# where strongly upwinded power-law faces leave couplings of a few 1e-40 (FP32
# subnormals), and natural convection at Ra 8e5 around a cylinder
VANISHING = {18: 400, 234: 370, 20: 730, 43: 1000, 118: 1100}


@pytest.mark.parametrize("precision", ["fp32_mixed", "fp32_refined", "fp32_stochastic"])
def test_fp32_paths_stay_finite_on_vanishing_couplings(precision):
    from ixchel2d_amd.simple import Cavity
    cases = C.batch_cases(256, seed=0)
    for k, steps in VANISHING.items():
        case = cases[k]
        g = Cavity(case, solver="pcr", precision=precision)
        g.run(steps + 50, cap=1)
        g64 = Cavity(case, solver="pcr", precision="fp64")
        g64.run(steps + 50, cap=1)
        for f in "uvtp":
            a = g.get(f)
            assert np.all(np.isfinite(a)), f"{case.name}: {f}"
        for f in "uvt":
            ref = g64.get(f)
            if np.max(np.abs(ref)) > 0.0:
                e = relerr(g.get(f), ref)
                print(f"[measure] {precision} {case.name} {f} {e:.2e}", flush=True)
                assert e <= VANISHING_TOL[precision], f"{case.name}: {f} rel err {e:.2e}"
        g.close()
        g64.close()


# measured on MI355X, largest over the cavities and fields: FP32-mixed 2.7e-4,
# FP32-refined 6.4e-6, FP32-stochastic 7.1e-6 (all at Re 3134, 1150 steps)
VANISHING_TOL = {"fp32_mixed": 1e-3, "fp32_refined": 3e-5, "fp32_stochastic": 3e-5}
