"""FP32-stochastic path (precision IXH_FP32_STOCHASTIC): the FP32 storage of
the FP32-mixed path (u, v, T, p', a_u, a_v in FP32; p, b_o, the norms, the u
residual and the Nusselt integrals in FP64) with the increment sweeps of the
FP32-refined path, whose row residuals are formed in FP64 from the FP32
fields; every new u, v, T, of the sweeps and of the correction, is
phi + delta formed in FP64 and stored with stochastic rounding.

The pin: from the same state (FP32-representable fields), each phase of this
path must give, node by node, one of the two FP32 neighbours of what the
FP32-refined path (the same FP32 coefficients and solve, FP64 storage)
gives, and the up/down choices must follow the position between them, which
round-to-nearest would not; every other output of the phase (p', b_o, p,
a_u, a_v, Res_u) must be equal bit for bit.  The FP32-refined path is
pinned against the FP64 path and the upstream fixtures by its own tests.

Measured on MI355X (bounds a few times these):
  - phase by phase: 1570 (nc18x14), 2264 (mixed22x17) and 8481 (nc34)
    rounded values, mean of (rounded up - position) -0.007 .. 0.000; of
    those below the midpoint 167, 251 and 1012 rounded up against 173, 249
    and 1037 expected;
  - upstream fixtures: nc18x14 1.3e-7, mixed22x17 1.8e-7, lid66 1.1e-7,
    nc34 4.0e-7 relative to each field's maximum;
  - 34x34 natural convection run toward the steady state: residual 9.8e-9
    after 10500 iterations (FP32-mixed stalls at 3e-8), hot-wall Nu 4.5e-7
    from the FP64 steady state;
  - 301x701, 5000 iterations: hot-wall Nu 8.4e-7 from FP64 (FP32-mixed
    4.4e-4), u, v, T within 1.9e-7 .. 2.8e-7 of their maxima, and within
    1.7e-7 .. 2.8e-7 of the upstream fixture;
  - immersed cylinder, 30 iterations: u, v, T within 1.8e-7 .. 2.3e-7 of
    the FP64 path."""
import numpy as np
import pytest

import cases as K
from ixchel2d_amd import case as C

pytestmark = pytest.mark.gpu

PREC = "fp32_stochastic"


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), 1e-300))


def _cavity(case, precision=PREC):
    from ixchel2d_amd.simple import Cavity
    return Cavity(case, solver="pcr", precision=precision)


def _f32_neighbours(x):
    """The FP32 values just below and just above each FP64 value (equal where
    x is representable)."""
    f = x.astype(np.float32)
    fd = f.astype(np.float64)
    lo = np.where(fd > x, np.nextafter(f, np.float32(-np.inf)), f).astype(np.float64)
    hi = np.where(fd < x, np.nextafter(f, np.float32(np.inf)), f).astype(np.float64)
    return lo, hi


PHASES = [
    ("momentum_u_y", lambda g: g.momentum_u(0), ["u"], ["au"]),
    ("momentum_u_x", lambda g: g.momentum_u(1), ["u"], ["au"]),
    ("momentum_v_y", lambda g: g.momentum_v(0), ["v"], ["av"]),
    ("momentum_v_x", lambda g: g.momentum_v(1), ["v"], ["av"]),
    ("pcorr_zero", lambda g: g.pcorr_zero(), [], []),
    ("pcorr_y", lambda g: g.pcorr(0), [], ["pcorr"]),
    ("pcorr_x", lambda g: g.pcorr(1), [], ["pcorr", "bo"]),
    ("correct", lambda g: g.correct(), ["u", "v"], ["p"]),
    ("energy_y", lambda g: g.energy(0), ["t"], []),
    ("energy_x", lambda g: g.energy(1), ["t"], []),
    ("residual_u", lambda g: g.residual_u(), [], ["resu"]),
    ("lag", lambda g: g.lag(), [], ["u_ant", "v_ant", "t_ant"]),
]
STATE = ("u", "v", "t", "p", "u_ant", "v_ant", "t_ant", "au", "av")


@pytest.mark.parametrize("name,steps", [("nc18x14", 3), ("mixed22x17", 2), ("nc34", 40)])
def test_phase_by_phase_is_stochastic_rounding_of_refined(name, steps):
    """Each phase from one shared state: the rounded fields are FP32
    neighbours of the FP32-refined result, everything else equal to it."""
    from oracle.pyoracle import Oracle
    case = K.CASES[name]()
    o = Oracle(case)
    o.run(steps)
    g = _cavity(case)
    r = _cavity(case, "fp32_refined")
    for f in STATE:
        with np.errstate(over="ignore"):  # 1e40 boundary coefficients: infinite in FP32
            a = o.get(f).astype(np.float32).astype(np.float64)
        g.set(f, a)
        r.set(f, a)
    fracs, ups = [], []
    for pname, call, rounded, exact in PHASES:
        call(g)
        call(r)
        for f in rounded:
            got, ref = g.get(f), r.get(f)
            lo, hi = _f32_neighbours(ref)
            assert np.all((got == lo) | (got == hi)), f"{name}/{pname}: {f} not an FP32 neighbour"
            between = hi > lo
            fracs.append(((ref - lo) / np.where(between, hi - lo, 1.0))[between])
            ups.append((got == hi)[between])
        for f in exact:
            got, ref = g.get(f), r.get(f)
            if f in ("au", "av"):
                # a boundary face's coefficient: 1e40 in FP64 storage,
                # infinite in FP32 (main.f90:309-310)
                face = ref == 1e40
                assert np.all(got[face] == np.inf), f"{name}/{pname}: {f} boundary faces"
                got, ref = got[~face], ref[~face]
            assert np.array_equal(got, ref), f"{name}/{pname}: {f} differs"
        # resynchronise on this path's fields, as the next phase reads them
        for f in STATE + ("pcorr",):
            r.set(f, g.get(f))
    frac, up = np.concatenate(fracs), np.concatenate(ups).astype(np.float64)
    n = frac.size
    assert n >= 500, n
    # unbiased overall, and rounding up below the midpoint about as often as
    # the position says (round-to-nearest never would)
    assert abs(float(np.mean(up - frac))) <= 4.0 / np.sqrt(n), (n, float(np.mean(up - frac)))
    low = frac < 0.5
    expected, seen = float(np.sum(frac[low])), float(np.sum(up[low]))
    print(f"[measure] {name}: {n} rounded values, bias {float(np.mean(up - frac)):.3f}; "
          f"below the midpoint {seen:.0f} rounded up, {expected:.1f} expected", flush=True)
    assert expected >= 20.0, expected
    assert abs(seen - expected) <= 4.0 * np.sqrt(expected) + 3.0, (seen, expected)
    g.close()
    r.close()


TOL = {"nc18x14": 5e-7, "mixed22x17": 1e-6, "lid66": 5e-7, "nc34": 2e-6}


@pytest.mark.parametrize("name", list(TOL))
def test_against_upstream_fixture(name):
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
        err = relerr(a, gold[f])
        print(f"[measure] {name} {f} {err:.2e}", flush=True)
        assert err <= TOL[name], f"{name}: {f} rel err {err:.2e}"
    np.testing.assert_allclose(h["maxbo"][:10], gold["history"][:10, 0], rtol=1e-4)


@pytest.mark.parametrize("precision", ["fp32_mixed", "fp32_refined", "fp32_stochastic"])
@pytest.mark.parametrize("name", ["nc18x14", "nc34", "nc200x24"])
def test_fused_step_equals_kernel_by_kernel(name, precision):
    """The fused 6-launch step (whose sweeps without a residual or fold run
    one 64-thread wave per line) against the kernel-by-kernel step (256
    threads per line): the same fields and norms bit for bit on every FP32
    path.  nc200x24 has 201-node x-lines, 4 rows per lane."""
    if name == "nc200x24":
        case = C.natural_convection_case(200, 24, height=0.3, Ra=1e4, iters=12, dt=0.01, seed=2)
    else:
        case = K.CASES[name]()
    case.params.itermax = 12
    fused = _cavity(case, precision)
    hf = fused.run(12)
    case.params.iter_ecuaci_max, case.params.iter_simple_max = 3, 3
    seq = _cavity(case, precision)
    hs = seq.run(12)
    for f in "uvtp":
        assert np.array_equal(fused.get(f), seq.get(f)), f
    for k in ("err_v", "err_p", "maxbo", "err_t", "residuo"):
        assert np.array_equal(hf[k], hs[k]), k


@pytest.mark.parametrize("precision", ["fp32_mixed", "fp32_stochastic"])
@pytest.mark.parametrize("mi,nj", [(20, 1100), (1100, 20), (700, 30), (2047, 12), (12, 2047)])
def test_long_lines(precision, mi, nj):
    """Lines past 1024 nodes (the LDS PCR solver in FP32) up to the longest
    built (2048 nodes) and 701-node x-lines (the T and u x-sweeps at 256
    threads): 30 fused iterations against the FP64 path."""
    case = C.natural_convection_case(mi, nj, height=nj / mi, Ra=1e4, iters=30, dt=0.05, seed=5)
    g = _cavity(case, precision)
    g.run(30)
    g64 = _cavity(case, "fp64")
    g64.run(30)
    for f in "uvtp":
        a = g.get(f)
        assert np.all(np.isfinite(a)), f
        e = relerr(a, g64.get(f))
        print(f"[measure] {precision} {mi}x{nj} {f} {e:.2e}", flush=True)
        assert e <= 1e-4, f


def test_deterministic():
    """The rounding bits come from the values and their nodes: two runs from
    the same start give the same bits."""
    case = K.nc34()
    a, b = _cavity(case), _cavity(case)
    a.run(300)
    b.run(300)
    for f in "uvtp":
        assert np.array_equal(a.get(f), b.get(f)), f


def _run_to(g, tol, every, limit):
    it, r = 0, np.inf
    while it < limit:
        g.run(every - 1, cap=1)
        h = g.run(1)
        it += every
        r = float(h["residuo"][-1])
        if r <= tol:
            break
    return it, r


def test_small_steady_case_reaches_residual_1e8():
    """34x34 natural convection (Ra 1e5, dt 0.01) to the FP64-measured
    residual 1e-8; the hot-wall Nusselt number and the fields against the
    FP64 path's steady state."""
    case = K.nc34()
    g = _cavity(case)
    it, r = _run_to(g, 1e-8, 500, 60000)
    print(f"[measure] nc34 steady: residual {r:.2e} after {it} iterations", flush=True)
    assert r <= 1e-8, (it, r)
    g64 = _cavity(case, "fp64")
    _run_to(g64, 1e-13, 5000, 60000)
    nu, nu64 = g.nusselt_x()[0], g64.nusselt_x()[0]
    err = abs(nu - nu64) / abs(nu64)
    print(f"[measure] nc34 steady: Nu err {err:.2e}", flush=True)
    assert err <= 2e-6
    for f in "uvt":
        assert relerr(g.get(f), g64.get(f)) <= 5e-6, f


def test_full_size_5000_iterations():
    case = K.nc301x701()
    gold = K.load_fixture("nc301x701")
    g = _cavity(case)
    g.run(5000)
    g64 = _cavity(case, "fp64")
    g64.run(5000)
    nu, nu64 = g.nusselt_x()[0], g64.nusselt_x()[0]
    err = abs(nu - nu64) / abs(nu64)
    print(f"[measure] nc301x701 5000: Nu err {err:.2e}", flush=True)
    assert err <= 5e-6
    for f in "uvt":
        e = relerr(g.get(f), g64.get(f))
        print(f"[measure] nc301x701 5000: {f} {e:.2e}", flush=True)
        assert e <= 1e-6, f
    for f in "uvtp":
        e = relerr(K.subsample(g.get(f)), gold[f + "_sub"])
        print(f"[measure] nc301x701 5000 vs upstream: {f} {e:.2e}", flush=True)
        assert e <= 1e-6, f


def test_immersed_cylinder():
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
    assert np.max(np.abs(g.get("t")[inside] - g64.get("t")[inside])) <= 1e-6
    for f in "uvt":
        e = relerr(g.get(f), g64.get(f))
        print(f"[measure] cylinder {f} {e:.2e}", flush=True)
        assert e <= 1e-6, f
    p32, p64 = g.get("p"), g64.get("p")
    assert relerr(p32[~inside], p64[~inside]) <= 1e-6


def test_batch_matches_single_cavity():
    from ixchel2d_amd.simple import Batch
    cases = [C.lid_driven_case(30, Re=400.0, iters=10), K.nc34(),
             C.natural_convection_case(26, 40, height=1.5, Ra=1e5, iters=10, dt=0.01, seed=4)]
    C.cylinder_sources(cases[0], 0.5, 0.5, 0.12, t_body=1.0)
    for shape in ((64, 4, False), (64, 1, True), (256, 1, False)):
        cav = [_cavity(c) for c in cases]
        b = Batch(cav, order="xcd")
        b.set_shape(*shape)
        b.run(12)
        for c, case in enumerate(cases):
            alone = _cavity(case)
            alone.run(12)
            for f in "uvtp":
                assert np.array_equal(cav[c].get(f), alone.get(f)), f"{shape} cavity {c}: {f}"
        b.close()
    with pytest.raises(Exception):
        Batch([_cavity(cases[0]), _cavity(cases[1], "fp32_mixed")])


def test_rejects_other_solvers_and_the_adjoint():
    from ixchel2d_amd import _lib
    from ixchel2d_amd.simple import Cavity
    with pytest.raises(_lib.IxhError):
        Cavity(K.nc18x14(), solver="thomas", precision=PREC)
    g = _cavity(K.nc18x14())
    with pytest.raises(_lib.IxhError):
        g.adjoint_record(4)
