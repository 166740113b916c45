"""The batch workload at its own size: 256 independent cavities of 65^2,
129^2, 257^2 and 129 x 301 cells (seed 0), half lid-driven (Re 100..3200),
half natural convection (Ra 1e3..1e6, Pr 0.71), 2 of every 16 with an
immersed cylinder held at T = 1, advanced together on the FP32-mixed path
for 50 steps, with the schedule (ordering, threads per line, residual
placement) the autotuner picks on a copy of the batch and each step replayed
as a captured HIP graph.

  - every cavity ends with the fields of the same cavity run alone through
    ixh_simple_iterations, bit for bit;
  - the hot-wall Nusselt number of every cavity with heat flow is within
    1e-4 (relative) of the same batch run on the FP64 path, on the
    FP32-mixed path (measured on MI355X after these 50 steps: largest 2.0e-5,
    median 3.9e-6), the FP32-refined path (largest 6.4e-8) and the
    FP32-stochastic path;
  - the 16 smallest cavities (65^2) agree with the FP64 oracle to the FP32
    bound 2e-5 relative (pressure compared outside the bodies: inside, it is
    rounding noise amplified by the 1e51 penalty)."""
import numpy as np
import pytest

from ixchel2d_amd import case as C
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

N = 256
STEPS = 50
TRIAL = 1
NU_TOL = 1e-4
ORACLE_TOL = 2e-5


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)


@pytest.fixture(scope="module")
def batch256():
    from ixchel2d_amd.simple import Batch, Cavity
    cases = C.batch_cases(N, seed=0)
    # the autotuner's trial steps are real time steps (trial + 4 per
    # schedule): tune on a copy of the batch, then run the tested one from
    # the start under the chosen schedule
    tune = [Cavity(c, solver="pcr", precision="fp32_mixed") for c in cases]
    bt = Batch(tune)
    sched, table = bt.autotune(trial_steps=TRIAL)
    bt.close()
    for c in tune:
        c.close()
    cav = [Cavity(c, solver="pcr", precision="fp32_mixed") for c in cases]
    b = Batch(cav, order=sched["order"], graph=True)
    b.set_shape(sched["line_threads"], sched["lines_per_group"], sched["residual_with_t"])
    hist = b.run(STEPS)
    print(f"\n[batch256] autotuned {sched}; timings {table}", flush=True)
    yield cases, cav, hist, STEPS
    b.close()
    for c in cav:
        c.close()


def test_batch256_equals_each_cavity_alone(batch256):
    from ixchel2d_amd.simple import Cavity
    cases, cav, hist, total = batch256
    for c, case in enumerate(cases):
        g = Cavity(case, solver="pcr", precision="fp32_mixed")
        h = g.run(total, cap=total)
        for f in "uvtp":
            assert np.array_equal(cav[c].get(f), g.get(f)), f"cavity {c} ({case.name}): {f}"
        assert np.array_equal(hist[c]["maxbo"], h["maxbo"][-STEPS:]), f"cavity {c}: maxbo"
        g.close()


def _nu_errors(cases, cav, ref):
    errs = []
    for c in range(len(cases)):
        nu64 = ref[c].nusselt_x()[0]
        nu32 = cav[c].nusselt_x()[0]
        if abs(nu64) < 1e-12:  # lid-driven cavity without a body: T = 0, no heat flow
            assert abs(nu32) < 1e-12, (c, nu32)
            continue
        errs.append((abs(nu32 - nu64) / abs(nu64), c))
    errs.sort()
    return errs


def _run_batch(cases, precision, total):
    from ixchel2d_amd.simple import Batch, Cavity
    cav = [Cavity(c, solver="pcr", precision=precision) for c in cases]
    b = Batch(cav, order="xcd")
    b.run(total, cap=1)
    b.close()
    return cav


def test_batch256_nusselt_against_fp64(batch256):
    cases, cav, _, total = batch256
    ref = _run_batch(cases, "fp64", total)
    refined = _run_batch(cases, "fp32_refined", total)
    stochastic = _run_batch(cases, "fp32_stochastic", total)
    for name, got, tol in (("fp32_mixed", cav, NU_TOL), ("fp32_refined", refined, NU_TOL),
                           ("fp32_stochastic", stochastic, NU_TOL)):
        errs = _nu_errors(cases, got, ref)
        e = np.array([x for x, _ in errs])
        print(f"\n[batch256] {name}: hot-wall Nu vs FP64 over {len(e)} cavities: max {e.max():.2e} "
              f"(cavity {errs[-1][1]}), median {np.median(e):.2e}", flush=True)
        assert len(e) >= N // 2
        assert e.max() <= tol, (name, errs[-5:])
    for r in ref + refined + stochastic:
        r.close()


def test_batch256_smallest_against_oracle(batch256):
    cases, cav, _, total = batch256
    small = sorted(range(N), key=lambda c: (cases[c].mesh.mi * cases[c].mesh.nj, c))[:16]
    for c in small:
        case = cases[c]
        o = Oracle(case)
        o.run(total)
        outside = np.ones((case.mesh.mi + 1, case.mesh.nj + 1), dtype=bool)
        if "t" in case.sources:
            outside = case.sources["t"][1] == 0.0
        for f in "uvtp":
            ref, got = o.get(f), cav[c].get(f)
            if f == "p":
                ref, got = ref[outside], got[outside]
            scale = float(np.max(np.abs(ref)))
            if scale == 0.0:
                assert float(np.max(np.abs(got))) == 0.0, (c, f)
                continue
            err = float(np.max(np.abs(got - ref))) / scale
            assert err <= ORACLE_TOL, f"cavity {c} ({case.name}): {f} rel err {err:.2e}"
