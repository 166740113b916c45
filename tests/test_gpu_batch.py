"""GPU parity of the batched path (ixh_batch_*): every cavity of a batch must
end with the fields and norms of the same cavity run alone through
ixh_simple_iterations (bit for bit: per cavity the batch runs the same kernel
bodies in the same order), and with the Thomas solver bit-identical to the
oracle.  Covers every workgroup ordering, graph replay, consecutive calls
(argument caches rebuilt per call), immersed cylinders, buoyancy in the v
equation (natural and mixed convection), both launch shapes (threads per
line, the u residual beside the T x-sweep or on its own) and the
autotuner."""
import numpy as np
import pytest

import cases as K
from ixchel2d_amd import case as C
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)


def _small_batch():
    cs = [
        C.lid_driven_case(18, Re=400.0, iters=20),
        C.natural_convection_case(20, 27, height=1.4, Ra=2e4, iters=20, dt=0.01, seed=1, mesh_kind="unif"),
        K.mixed22x17(),
        C.natural_convection_case(34, 50, height=1.5, Ra=1e5, iters=20, dt=0.01, seed=2, beta=1.1),
        C.lid_driven_case(40, Re=1000.0, iters=20),
        C.natural_convection_case(26, 26, Ra=1e5, iters=20, dt=0.01, seed=3, mesh_kind="unif"),
    ]
    C.cylinder_sources(cs[4], 0.4, 0.55, 0.12, t_body=1.0)
    C.cylinder_sources(cs[5], 0.6, 0.4, 0.1, t_body=1.0)
    return cs


def _run_alone(case, solver, steps):
    from ixchel2d_amd.simple import Cavity
    g = Cavity(case, solver=solver)
    h = g.run(steps)
    return {f: g.get(f) for f in "uvtp"}, h


HIST = ("err_v", "err_p", "maxbo", "err_t", "residuo")


@pytest.mark.parametrize("order,graph", [("cavity", False), ("largest", False), ("xcd", False), ("xcd", True)])
def test_batch_matches_single_cavity_and_oracle_thomas(order, graph):
    from ixchel2d_amd.simple import Batch, Cavity
    cases = _small_batch()
    cav = [Cavity(c, solver="thomas") for c in cases]
    b = Batch(cav, order=order, graph=graph)
    h1 = b.run(7)
    h2 = b.run(5)
    for c, case in enumerate(cases):
        o = Oracle(case)
        ho = o.run(12)
        for f in "uvtp":
            assert np.array_equal(cav[c].get(f), o.get(f)), f"cavity {c} ({case.name}): {f} differs from the oracle"
        for k in HIST:
            got = np.r_[h1[c][k], h2[c][k]]
            np.testing.assert_allclose(got, ho[k], rtol=1e-12, atol=1e-300, err_msg=f"cavity {c} {k}")
    b.close()


@pytest.mark.parametrize("solver", ["pcr", "pcr_exact"])
def test_batch_matches_single_cavity_pcr(solver):
    from ixchel2d_amd.simple import Batch, Cavity
    cases = _small_batch()
    cav = [Cavity(c, solver=solver) for c in cases]
    b = Batch(cav, order="xcd", graph=True)
    hb = b.run(15)
    for c, case in enumerate(cases):
        alone, ha = _run_alone(case, solver, 15)
        for f in "uvtp":
            assert np.array_equal(cav[c].get(f), alone[f]), f"cavity {c}: {f}"
        for k in HIST:
            assert np.array_equal(hb[c][k], ha[k]), f"cavity {c}: {k}"
        o = Oracle(case)
        o.run(15)
        # inside a penalised body the pressure is rounding noise amplified
        # by the 1e51 penalty (about 1e22 there against 0.2 outside); only
        # the reference's own rounding (the Thomas path) reproduces it, so
        # the PCR paths are compared on the p-nodes outside the body
        outside = np.ones((case.mesh.mi + 1, case.mesh.nj + 1), dtype=bool)
        if "t" in case.sources:
            outside = case.sources["t"][1] == 0.0
        for f in "uvtp":
            ref, got = o.get(f), cav[c].get(f)
            if f == "p":
                ref, got = ref[outside], got[outside]
            err = np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300)
            assert err <= 1e-10, f"cavity {c}: {f} rel err {err:.2e} vs oracle"
    b.close()


def test_batch_autotune_then_run():
    from ixchel2d_amd.simple import Batch, Cavity
    cases = _small_batch()
    cav = [Cavity(c, solver="pcr") for c in cases]
    b = Batch(cav)
    sched, table = b.autotune(trial_steps=1)
    # every line is short and no u equation reads T: 3 orderings x graph or
    # not x 5 shapes (64 threads: one line, or a group of lines; the residual
    # beside T with one line; 256 threads, the residual either way)
    assert len(table) == 30 and all(v > 0 for v in table.values())
    assert {k[0] for k in table} == {"cavity", "largest", "xcd"}
    assert {k[2:] for k in table} == {(64, 1, False), (64, 1, True), (64, 4, False), (256, 1, False),
                                      (256, 1, True)}
    best = (sched["order"], sched["graph"], sched["line_threads"], sched["lines_per_group"],
            sched["residual_with_t"])
    assert table[best] == min(table.values())
    b.run(4)
    steps = len(table) * (1 + 4) + 4
    for c, case in enumerate(cases):
        alone, _ = _run_alone(case, "pcr", steps)
        for f in "uvtp":
            assert np.array_equal(cav[c].get(f), alone[f]), f"cavity {c}: {f}"
    b.close()


@pytest.mark.parametrize("solver,precision,line_threads,lines_per_group,residual_with_t",
                         [("pcr", "fp64", 64, 1, False), ("pcr", "fp64", 64, 1, True), ("pcr", "fp64", 64, 4, False),
                          ("pcr", "fp64", 256, 1, False), ("pcr", "fp64", 256, 1, True),
                          ("thomas", "fp64", 256, 1, True), ("thomas", "fp64", 256, 1, False),
                          ("pcr", "fp32_mixed", 64, 4, False), ("pcr", "fp32_refined", 64, 4, False)])
def test_batch_shapes_match_single_cavity(solver, precision, line_threads, lines_per_group, residual_with_t):
    """Threads per line, lines per workgroup and the residual's placement
    change the launch shape only: every cavity still equals its run alone,
    bit for bit, over calls that rebuild the arguments, and a cavity whose u
    equation reads T (Ri_x != 0) sends the residual back to a launch of its
    own."""
    from ixchel2d_amd.simple import Batch, Cavity
    cases = _small_batch()
    cav = [Cavity(c, solver=solver, precision=precision) for c in cases]
    b = Batch(cav, order="xcd")
    b.set_shape(line_threads, lines_per_group, residual_with_t)
    hb = b.run(5)
    alone = [Cavity(c, solver=solver, precision=precision) for c in cases]
    ha = [a.run(5) for a in alone]
    for c in range(len(cases)):
        for f in ("u", "v", "t", "p", "resu"):
            assert np.array_equal(cav[c].get(f), alone[c].get(f)), f"cavity {c}: {f}"
        for k in HIST:
            assert np.array_equal(hb[c][k], ha[c][k]), f"cavity {c}: {k}"
    # buoyancy in the u equation of one cavity, then 3 more steps
    for x in (cav[3], alone[3]):
        x.set_params(ri_x=0.4)
    hb = b.run(3)
    for c in range(len(cases)):
        ha = alone[c].run(3)
        for f in ("u", "v", "t", "p", "resu"):
            assert np.array_equal(cav[c].get(f), alone[c].get(f)), f"cavity {c} after ri_x: {f}"
        assert np.array_equal(hb[c]["residuo"], ha["residuo"]), f"cavity {c}: residuo"
    b.close()
    for x in cav + alone:
        x.close()


def test_batch_shape_rejects_one_wave_for_thomas():
    from ixchel2d_amd import _lib
    from ixchel2d_amd.simple import Batch, Cavity
    cav = [Cavity(c, solver="thomas") for c in _small_batch()[:2]]
    b = Batch(cav)
    with pytest.raises(_lib.IxhError, match="one wave per line"):
        b.set_shape(64, 1, True)
    with pytest.raises(_lib.IxhError, match="threads per line"):
        b.set_shape(128, 1, True)
    with pytest.raises(_lib.IxhError, match="one-wave lines"):
        b.set_shape(256, 4, False)
    with pytest.raises(_lib.IxhError, match="lines per group"):
        b.set_shape(256, 3, False)
    b.close()
    for x in cav:
        x.close()


def test_mixed_size_batch_of_32():
    """32 cavities of the mixed-size batch workload (65^2 .. 257^2 and
    129 x 301 cells, lid-driven and natural convection, cylinders)."""
    from ixchel2d_amd.simple import Batch, Cavity
    cases = C.batch_cases(32, seed=7)
    cav = [Cavity(c, solver="pcr") for c in cases]
    b = Batch(cav, order="xcd")
    hb = b.run(6)
    for c, case in enumerate(cases):
        alone, ha = _run_alone(case, "pcr", 6)
        for f in "uvtp":
            assert np.array_equal(cav[c].get(f), alone[f]), f"cavity {c} ({case.name}): {f}"
        assert np.array_equal(hb[c]["maxbo"], ha["maxbo"])
    b.close()


def test_batch_rejects_multi_pass_loops_and_mixed_solvers():
    from ixchel2d_amd import _lib
    from ixchel2d_amd.simple import Batch, Cavity
    cases = _small_batch()[:2]
    cav = [Cavity(cases[0], solver="pcr"), Cavity(cases[1], solver="thomas")]
    with pytest.raises(_lib.IxhError):
        Batch(cav)
    ctrl = Cavity(K.ctrl18x14(), solver="pcr")
    b = Batch([ctrl])
    with pytest.raises(_lib.IxhError):
        b.run(2)
    b.close()
