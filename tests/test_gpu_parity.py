"""GPU parity: the HIP path through the C-ABI (libixchel2d.so) against the
CPU oracle and the upstream golden fixtures.

Tolerances: with the Thomas line solver the device arithmetic follows
upstream operation by operation (no FMA contraction, IEEE division), so
fields must be bit-identical.  "pcr_exact" keeps upstream's assembly and
changes only the line-solve rounding (1e-12 per phase); "pcr" (the default
path) also uses the division-free assembly (1e-11 per phase).  Whole runs
must agree with upstream to 1e-10 relative (the FP64 target), norms to 1e-9.
"""
import numpy as np
import pytest

import cases as K
from ixchel2d_amd import case as C
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

REL = 1e-10


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)


def _cavity(case, solver):
    from ixchel2d_amd.simple import Cavity
    return Cavity(case, solver=solver)


def relerr(a, b):
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale


PHASES = [
    ("momentum_u_y", lambda g: g.momentum_u(0), ["u", "au"]),
    ("momentum_u_x", lambda g: g.momentum_u(1), ["u", "au"]),
    ("momentum_v_y", lambda g: g.momentum_v(0), ["v", "av"]),
    ("momentum_v_x", lambda g: g.momentum_v(1), ["v", "av"]),
    ("pcorr_zero", lambda g: g.pcorr_zero(), []),
    ("pcorr_y", lambda g: g.pcorr(0), ["pcorr"]),
    ("pcorr_x", lambda g: g.pcorr(1), ["pcorr", "bo"]),
    ("correct", lambda g: g.correct(), ["u", "v", "p"]),
    ("energy_y", lambda g: g.energy(0), ["t"]),
    ("energy_x", lambda g: g.energy(1), ["t"]),
    ("residual_u", lambda g: g.residual_u(), ["resu"]),
    ("lag", lambda g: g.lag(), ["u_ant", "v_ant", "t_ant"]),
]


def _warm_states(case, steps):
    """Oracle and GPU states with the same non-trivial fields: the oracle
    runs `steps` iterations and the GPU state receives its fields."""
    o = Oracle(case)
    if steps:
        o.run(steps)
    return o


PHASE_TOL = {"pcr_exact": 1e-12, "pcr": 1e-11}


@pytest.mark.parametrize("name,steps", [("nc18x14", 3), ("mixed22x17", 2)])
@pytest.mark.parametrize("solver", ["thomas", "pcr_exact", "pcr"])
def test_phase_by_phase(name, steps, solver):
    case = K.CASES[name]()
    o = _warm_states(case, steps)
    g = _cavity(case, solver)
    for f in ("u", "v", "t", "p", "u_ant", "v_ant", "t_ant", "au", "av"):
        g.set(f, o.get(f))
    for pname, call, outs in PHASES:
        o.phase(pname if pname != "residual_u" else "residual_u")
        call(g)
        for f in outs:
            a, b = g.get(f), o.get(f)
            if solver == "thomas":
                assert np.array_equal(a, b), f"{name}/{pname}: {f} not bit-identical ({relerr(a, b):.2e})"
            else:
                tol = PHASE_TOL[solver]
                assert relerr(a, b) <= tol, f"{name}/{pname}: {f} rel err {relerr(a, b):.2e}"
            g.set(f, b)  # resynchronise so each phase is judged on its own


@pytest.mark.parametrize("name", ["nc18x14", "mixed22x17", "ctrl18x14", "lid66", "nc34"])
@pytest.mark.parametrize("solver", ["thomas", "pcr_exact", "pcr"])
def test_run_against_upstream_fixture(name, solver):
    case = K.CASES[name]()
    gold = K.load_fixture(name)
    g = _cavity(case, solver)
    h = g.run(case.params.itermax)
    for f in "uvtp":
        a = g.get(f)
        if solver == "thomas":
            assert np.array_equal(a, gold[f]), f"{name}: {f} differs from upstream ({relerr(a, gold[f]):.2e})"
        else:
            assert relerr(a, gold[f]) <= REL, f"{name}: {f} rel err {relerr(a, gold[f]):.2e}"
    logged = h["logged"].astype(bool)
    hist = np.c_[h["maxbo"], h["residuo"]][logged]
    assert hist.shape == gold["history"].shape, "SIMPLE iteration count differs from upstream"
    np.testing.assert_allclose(hist, gold["history"], rtol=1e-9, atol=1e-15)


@pytest.mark.parametrize("mi,nj", [(20, 900), (20, 1100), (1100, 20), (700, 30), (1800, 12), (12, 1800),
                                   (2047, 12), (12, 2047)])
def test_long_lines_against_oracle(mi, nj):
    """Line lengths the 301x701 run does not reach, each through its own
    launch shapes: 901-node y-lines (4 rows per thread), 1101-node y- and
    x-lines (256 threads, the LDS PCR solver for lines past 1024 nodes),
    701-node x-lines (the T/u x launch with 256 threads), 1801-node x- and
    y-lines (8 rows per thread; the p' x-sweep with the correction then
    holds two lines in 135 KB of LDS), and 2048-node x- and y-lines, the
    longest the library builds.  30 fused iterations against the oracle."""
    case = C.natural_convection_case(mi, nj, height=nj / mi, Ra=1e4, iters=30, dt=0.05, seed=5)
    case.params.paq_itera = 10
    g = _cavity(case, "pcr")
    g.run(30)
    o = Oracle(case)
    o.run(30)
    for f in "uvtp":
        assert relerr(g.get(f), o.get(f)) <= REL, f"{f}: rel err {relerr(g.get(f), o.get(f)):.2e}"


@pytest.mark.parametrize("mi,nj", [(4, 4), (4, 30), (30, 4), (5, 7)])
@pytest.mark.parametrize("solver", ["thomas", "pcr"])
def test_smallest_meshes_against_oracle(mi, nj, solver):
    """The smallest meshes the library takes (mi, nj >= 4: u x-lines of 4
    nodes, v y-lines of 4): 30 iterations of the fused step against the
    oracle, bit for bit with upstream's Thomas order."""
    case = C.natural_convection_case(mi, nj, height=nj / mi, Ra=1e4, iters=30, dt=0.05, seed=5)
    case.params.paq_itera = 10
    g = _cavity(case, solver)
    g.run(30)
    o = Oracle(case)
    o.run(30)
    for f in "uvtp":
        if solver == "thomas":
            assert np.array_equal(g.get(f), o.get(f)), f
        else:
            assert relerr(g.get(f), o.get(f)) <= REL, f"{f}: rel err {relerr(g.get(f), o.get(f)):.2e}"


@pytest.mark.parametrize("mi,nj", [(2048, 12), (12, 2048)])
def test_lines_past_the_longest_are_rejected(mi, nj):
    """A mesh whose lines would have 2049 nodes is refused at creation with
    the library's error, not run."""
    from ixchel2d_amd import _lib
    case = C.natural_convection_case(mi, nj, height=nj / mi, Ra=1e4, iters=1, dt=0.05, seed=5)
    with pytest.raises(_lib.IxhError, match="lines longer than 2048"):
        _cavity(case, "pcr")


@pytest.mark.parametrize("precision,tol", [("fp32_mixed", 5e-5), ("fp32_refined", 5e-6)])
@pytest.mark.parametrize("mi,nj", [(20, 900), (20, 1100), (1100, 20), (700, 30), (2047, 12), (12, 2047)])
def test_long_lines_fp32_paths(mi, nj, precision, tol):
    """The FP32 paths through the same long-line launch shapes (FP32 wave
    solver up to 1024 nodes, FP32 LDS PCR beyond), 30 fused iterations
    against the FP64 oracle.  Measured: FP32-mixed 2e-6..6e-6, FP32-refined
    1e-8..5e-7 of each field's maximum."""
    case = C.natural_convection_case(mi, nj, height=nj / mi, Ra=1e4, iters=30, dt=0.05, seed=5)
    case.params.paq_itera = 10
    from ixchel2d_amd.simple import Cavity
    g = Cavity(case, solver="pcr", precision=precision)
    g.run(30)
    o = Oracle(case)
    o.run(30)
    for f in "uvtp":
        a = g.get(f)
        assert np.all(np.isfinite(a)), f
        assert relerr(a, o.get(f)) <= tol, f"{f}: rel err {relerr(a, o.get(f)):.2e}"


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed", "fp32_refined"])
def test_fused_steps_split_across_calls(precision):
    """A fused step runs the residual of the step before it beside its v
    y-sweep, folds the norms beside its v x-sweep and alternates the v norm
    slots; runs split over several calls (5 + 1 + 6 steps) must give the
    fields and the history of one 12-step run bit for bit."""
    from ixchel2d_amd.simple import Cavity
    case = K.mixed22x17()
    whole = Cavity(case, solver="pcr", precision=precision)
    hw = whole.run(12)
    split = Cavity(case, solver="pcr", precision=precision)
    parts = [split.run(n) for n in (5, 1, 6)]
    for f in "uvtp":
        assert np.array_equal(split.get(f), whole.get(f)), f
    for k in ("err_v", "err_p", "maxbo", "err_t", "residuo"):
        assert np.array_equal(np.concatenate([h[k] for h in parts]), hw[k]), k


@pytest.mark.parametrize("name", ["nc18x14", "mixed22x17", "nc34"])
def test_fused_step_equals_kernel_by_kernel(name):
    """With convergence thresholds of 1e30 every loop stops after one pass
    whatever its pass limit; non-negative limits make the driver run the
    kernels one after the other (upstream order, norms read back after each
    pass), negative ones select the fused 6-launch step.  Both must give the
    same fields and norms bit for bit."""
    case = K.CASES[name]()
    case.params.itermax = 12
    fused = _cavity(case, "pcr")
    hf = fused.run(12)
    case.params.iter_ecuaci_max, case.params.iter_simple_max = 3, 3
    seq = _cavity(case, "pcr")
    hs = seq.run(12)
    for f in "uvtp":
        assert np.array_equal(fused.get(f), seq.get(f)), f
    for k in ("err_v", "err_p", "maxbo", "err_t", "residuo"):
        assert np.array_equal(hf[k], hs[k]), k


def test_control_flow_pass_counts_match_oracle():
    case = K.ctrl18x14()
    g = _cavity(case, "thomas")
    o = Oracle(case)
    hg = g.run(case.params.itermax)
    ho = o.run(case.params.itermax)
    for k in ("passes_mom", "passes_p", "passes_t", "simple_iter", "logged"):
        assert np.array_equal(hg[k], ho[k]), k
    for k in ("err_v", "err_p", "maxbo", "err_t", "residuo"):
        np.testing.assert_allclose(hg[k], ho[k], rtol=1e-9, atol=1e-15)


def test_nusselt_against_oracle_and_upstream():
    case = K.nc34()
    gold = K.load_fixture("nc34")
    g = _cavity(case, "thomas")
    g.run(1000)
    o = Oracle(case)
    o.run(1000)
    assert g.nusselt_y() == o.nusselt_y()
    assert g.nusselt_x() == o.nusselt_x()
    n0, n1 = g.nusselt_y()
    assert n0 == pytest.approx(gold["nusselt"][0, 1], rel=1e-14, abs=1e-17)
    assert -n1 == pytest.approx(gold["nusselt"][0, 2], rel=1e-14, abs=1e-17)


def test_immersed_cylinder_sources():
    case = C.natural_convection_case(40, 48, height=1.2, Ra=1e5, iters=30, dt=0.01, seed=5)
    C.cylinder_sources(case, 0.55, 0.6, 0.15, t_body=1.0)
    g = _cavity(case, "thomas")
    o = Oracle(case)
    g.run(30)
    o.run(30)
    for f in "uvtp":
        assert np.array_equal(g.get(f), o.get(f)), f
    # inside the body the penalty pins T to the body temperature and u, v
    # to the penalty value 1e-11
    m = case.mesh
    X, Y = np.meshgrid(m.xp, m.yp, indexing="ij")
    inside = (X - 0.55) ** 2 + (Y - 0.6) ** 2 <= 0.15 ** 2
    t = g.get("t")
    assert np.max(np.abs(t[inside] - 1.0)) < 1e-12
    mu = case.sources["u"][1] != 0.0
    mv = case.sources["v"][1] != 0.0
    assert mu.sum() > 10 and mv.sum() > 10
    assert np.max(np.abs(g.get("u")[mu])) <= 1e-10
    assert np.max(np.abs(g.get("v")[mv])) <= 1e-10


@pytest.mark.parametrize("name", ["ib161x41", "ib161x41t", "ib201x41c"])
@pytest.mark.parametrize("solver", ["thomas", "pcr"])
def test_immersed_bodies_against_upstream(solver, name):
    """Upstream's own immersed bodies: the reference built with
    definir_cuerpo(..., 'cuadr' | 'trian' | 'chime') (main.f90:327), 200
    iterations of a long lid-driven mixed-convection cavity.  Thomas: every
    field bit for bit, the penalty-amplified pressure inside the body
    included; PCR: 1e-10 relative, the pressure compared outside the body
    (inside it is rounding noise amplified by the 1e51 penalty).  u and v
    inside sit at the penalty value, about 1e-11; 'chime' holds T = 1."""
    body = K.REF_BODY[name]
    case = K.with_body(K.CASES[name]())
    gold = K.load_fixture(name)
    g = _cavity(case, solver)
    h = g.run(case.params.itermax)
    inside = C.upstream_body_mask(case.mesh, body)
    outside = ~inside
    for f in "uvtp":
        a, b = g.get(f), gold[f]
        if solver == "thomas":
            assert np.array_equal(a, b), f
        else:
            if f == "p":
                a, b = a[outside], b[outside]
            assert relerr(a, b) <= REL, f"{f}: {relerr(a, b):.2e}"
    mu = case.sources["u"][1] != 0.0
    mv = case.sources["v"][1] != 0.0
    assert np.max(np.abs(g.get("u")[mu])) <= 1e-10
    assert np.max(np.abs(g.get("v")[mv])) <= 1e-10
    if body == "chime":
        assert np.max(np.abs(g.get("t")[inside] - 1.0)) <= 1e-12
    logged = h["logged"].astype(bool)
    hist = np.c_[h["maxbo"], h["residuo"]][logged]
    assert hist.shape == gold["history"].shape
    np.testing.assert_allclose(hist[:, 0], gold["history"][:, 0], rtol=1e-9, atol=1e-15)
    # residuo here (1e22 to 1e29) is the u residual of the penalised rows:
    # the rounding of u inside the body times the 1e51 penalty, so the line
    # solver's rounding shows in it.  With PCR it follows upstream's to 3.6e-7
    # for the bodies on the wall, where u is small; the rod of 'chime' sits
    # in the moving fluid, where the rounding noise itself differs between
    # the solvers (up to 37%, at 1e22 against a physical residual below
    # 1e-2), so there only its magnitude is compared.
    if solver == "thomas":
        np.testing.assert_allclose(hist[:, 1], gold["history"][:, 1], rtol=1e-12, atol=1e-15)
    elif body != "chime":
        np.testing.assert_allclose(hist[:, 1], gold["history"][:, 1], rtol=1e-6, atol=1e-15)
    else:
        ratio = hist[:, 1] / gold["history"][:, 1]
        assert np.all((ratio > 0.5) & (ratio < 2.0)), ratio


@pytest.mark.parametrize("solver", ["thomas", "pcr"])
def test_tridiagonal_against_dense_solve(solver):
    from ixchel2d_amd.simple import tridiagonal
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 64, 65, 128, 129, 256, 257, 702, 1024, 1025, 2048):
        nsys = 5
        a = rng.uniform(-1, 0, (nsys, n))
        c = rng.uniform(-1, 0, (nsys, n))
        b = 2.2 + rng.uniform(0, 1, (nsys, n))
        r = rng.standard_normal((nsys, n))
        x, _ = tridiagonal(a, b, c, r, solver=solver)
        for k in range(nsys):
            if n == 1:
                np.testing.assert_allclose(x[k], r[k] / b[k], rtol=1e-15)
                continue
            A = np.diag(b[k]) + np.diag(a[k, 1:], -1) + np.diag(c[k, :-1], 1)
            np.testing.assert_allclose(x[k], np.linalg.solve(A, r[k]), rtol=1e-12, atol=1e-13)
        if solver == "thomas":
            for k in range(nsys):
                from oracle.pyoracle import tridiagonal as otri
                xo, bo = otri(a[k], b[k], c[k], r[k])
                assert np.array_equal(x[k], xo)


def test_tridiagonal_vanishing_couplings():
    """Sub- and super-diagonal entries of 1e-310 (subnormal) where the wave
    solver's chunks meet (512 rows: 8 per lane), as strongly upwinded faces
    give: the solve stays finite and equals a dense solve."""
    from ixchel2d_amd.simple import tridiagonal
    rng = np.random.default_rng(11)
    n, nsys = 512, 4
    a = rng.uniform(-1, 0, (nsys, n))
    c = rng.uniform(-1, 0, (nsys, n))
    b = 2.2 + rng.uniform(0, 1, (nsys, n))
    r = rng.standard_normal((nsys, n))
    a[:, 8::8] = -1e-310
    c[:, 7::8] = -1e-310
    a[1, 8::16] = 0.0  # exact zeros beside them
    c[2, 15::16] = 0.0
    x, _ = tridiagonal(a, b, c, r, solver="pcr")
    assert np.all(np.isfinite(x))
    for k in range(nsys):
        A = np.diag(b[k]) + np.diag(a[k, 1:], -1) + np.diag(c[k, :-1], 1)
        np.testing.assert_allclose(x[k], np.linalg.solve(A, r[k]), rtol=1e-12, atol=1e-13)


def test_full_size_5000_iterations_against_upstream():
    """301x701, 5000 SIMPLE iterations (tanh mesh, Ra=1e5): PCR path vs the
    upstream CPU build, on the stored subsample, centre lines and norms, the
    temperature next to every wall, upstream's nusselt_promedio_y record of
    every 1000 iterations (mod_postproceso.f90:100-215, main.f90:1226-1247)
    and the hot-wall mean Nusselt number of the reference's own field."""
    from ixchel2d_amd.adjoint import nusselt_x_weights
    case = K.nc301x701()
    gold = K.load_fixture("nc301x701")
    g = _cavity(case, "pcr")
    hs = []
    for k in range(5):
        hs.append(g.run(1000))
        # the horizontal walls are adiabatic, so these integrals are small
        # differences of near-equal temperatures (about 5e-4 against a
        # hot-wall Nu of 3): 1e-10 relative plus 1e-9 absolute
        n0, n1 = g.nusselt_y()
        assert n0 == pytest.approx(gold["nusselt"][k, 1], rel=1e-10, abs=1e-9), (k, n0)
        assert -n1 == pytest.approx(gold["nusselt"][k, 2], rel=1e-10, abs=1e-9), (k, n1)
    h = {f: np.concatenate([x[f] for x in hs]) for f in ("maxbo", "residuo")}
    for f in "uvtp":
        a = g.get(f)
        sub = K.subsample(a)
        assert relerr(sub, gold[f + "_sub"]) <= REL, f"{f}: {relerr(sub, gold[f + '_sub']):.2e}"
        assert relerr(a[a.shape[0] // 2, :], gold[f + "_xmid"]) <= REL
        assert relerr(a[:, a.shape[1] // 2], gold[f + "_ymid"]) <= REL
        assert float(np.sqrt(np.sum(a * a))) == pytest.approx(float(gold[f + "_l2"]), rel=REL)
    t = g.get("t")
    for key, part in (("t_cols_w", t[:3, :]), ("t_cols_e", t[-3:, :]), ("t_rows_s", t[:, :3]),
                      ("t_rows_n", t[:, -3:])):
        assert relerr(part, gold[key]) <= REL, f"{key}: {relerr(part, gold[key]):.2e}"
    # the vertical-wall mean Nusselt numbers of the reference's temperature
    # (weights of the same definition, pinned against the oracle on the CPU)
    tref = np.zeros_like(t)
    tref[:3, :] = gold["t_cols_w"]
    tref[-3:, :] = gold["t_cols_e"]
    nu_w_ref = float(np.sum(nusselt_x_weights(case.mesh, "west") * tref))
    nu_e_ref = float(np.sum(nusselt_x_weights(case.mesh, "east") * tref))
    nu_w, nu_e = g.nusselt_x()
    assert nu_w == pytest.approx(nu_w_ref, rel=REL), (nu_w, nu_w_ref)
    assert nu_e == pytest.approx(nu_e_ref, rel=REL), (nu_e, nu_e_ref)
    hist = np.c_[h["maxbo"], h["residuo"]]
    np.testing.assert_allclose(hist, gold["history"], rtol=1e-8, atol=1e-15)


def test_fused_step_launch_timing():
    """Profiling on: each of the 6 fused launches records its kernel's start
    and end; every launch kind is counted once per step."""
    import time
    import torch
    case = K.nc34()
    g = _cavity(case, "pcr")
    g.run(10)
    g.set_profiling(True)
    torch.cuda.synchronize()
    g.run(100)
    kt = g.kernel_times()
    g.set_profiling(False)
    names = ["multi_v_y+residual", "multi_v_x+fold", "step_sweep_p_y", "step_p_x+correct",
             "multi_t_y+u_y", "multi_t_x+u_x"]
    for n in names:
        ms, cnt = kt[n]
        assert cnt == 100 and ms > 0.0, (n, kt[n])
    # the profiled run gives the same fields as an unprofiled one
    h = _cavity(case, "pcr")
    h.run(110)
    for f in "uvtp":
        assert np.array_equal(g.get(f), h.get(f)), f


def test_state_on_a_torch_stream():
    """ixh_set_stream onto a torch stream (non-blocking with respect to the
    null stream): boundary conditions and sources changed between runs wait
    for the work queued on that stream, and the fields equal a run on the
    state's own stream bit for bit."""
    import torch
    case = C.natural_convection_case(34, 50, height=1.5, Ra=1e5, iters=20, dt=0.01, seed=3)
    C.cylinder_sources(case, 0.5, 0.7, 0.1, t_body=1.0)
    from ixchel2d_amd import _lib
    typ, val = case.bcs["t"].expand(case.mesh)["a"]
    hot = np.asarray(val) * 1.2

    def drive(c):
        # the T y-sweep is still queued when the boundary values change
        c.run(15)
        c.energy(_lib.DIR_Y)
        c.set_bc("t", "a", typ, hot)
        c.energy(_lib.DIR_X)
        c.run(15)

    ref = _cavity(case, "pcr")
    drive(ref)
    st = torch.cuda.Stream()
    g = _cavity(case, "pcr")
    g.set_stream(st.cuda_stream)
    drive(g)
    for f in "uvtp":
        assert np.array_equal(g.get(f), ref.get(f)), f
    g.set_stream(None)
    g.run(2)
    ref.run(2)
    for f in "uvtp":
        assert np.array_equal(g.get(f), ref.get(f)), f
