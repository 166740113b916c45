"""GPU checks of the discrete energy adjoint (ixh_adjoint_record /
ixh_adjoint_energy) on a flow that does not depend on T (lid-driven cavity
carrying a passive temperature), where the adjoint is the exact gradient of
the K-step map and J = Nu(T^K) is affine in the wall values:
  - against central finite differences on the GPU (h = 1e-6): 1e-6 relative;
  - against the FP64 reference gradient from the oracle (exact differences
    of the affine map on the CPU): 1e-8 relative;
  - dJ/dT(initial) against the change of J under a random perturbation of
    the initial temperature (dot-product test): 1e-8 relative;
  - checkpoint spacing (K not a multiple of it) does not change the result,
    and the state ends as after the recorded run."""
import numpy as np
import pytest

from ixchel2d_amd import adjoint as A
from ixchel2d_amd import case as C
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

K = 60


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)


@pytest.fixture(scope="module")
def problem():
    from ixchel2d_amd.simple import Cavity
    case = C.heated_lid_case(34, Re=100.0, dt=0.05)
    warm = Cavity(case, solver="pcr")
    warm.run(100)
    start = {f: warm.get(f) for f in ("u", "v", "t", "p")}
    return A.HotWallProblem(case, ncontrols=8, steps=K, every=20, start=start)


def _controls(n, seed=1):
    rng = np.random.default_rng(seed)
    return 1.0 + rng.uniform(-0.1, 0.1, n)


def test_adjoint_against_central_differences(problem):
    c = _controls(8)
    problem.forward(c)
    g = problem.gradient()
    h = 1e-6
    fd = np.zeros_like(c)
    for i in range(len(c)):
        e = np.zeros_like(c)
        e[i] = h
        fd[i] = (problem.forward(c + e) - problem.forward(c - e)) / (2 * h)
    assert np.max(np.abs(g - fd)) <= 1e-6 * np.max(np.abs(fd)), (g, fd)


def test_adjoint_against_oracle_reference_gradient(problem):
    c = _controls(8, seed=2)
    problem.forward(c)
    g = problem.gradient()
    case = problem.case

    def J(cc):
        o = Oracle(case)
        typ = problem.types_a
        o.set_bc("t", "a", typ, problem.wall_values(cc))
        for f, arr in problem.start.items():
            o.set(f, arr)
            if f in ("u", "v", "t"):
                o.set(f + "_ant", arr)
        o.run(K)
        return o.nusselt_x()[0]

    j0 = J(c)
    ref = np.array([J(c + np.eye(len(c))[i]) - j0 for i in range(len(c))])
    assert np.max(np.abs(g - ref)) <= 1e-8 * np.max(np.abs(ref)), (g, ref)


def test_initial_temperature_adjoint_dot_product(problem):
    c = _controls(8, seed=3)
    problem.forward(c)
    _, gt0 = problem.cav.adjoint_energy(problem.W)
    j0 = problem.cav.nusselt_x()[0]
    rng = np.random.default_rng(4)
    d = rng.standard_normal(problem.start["t"].shape) * 1e-3
    saved = problem.start["t"]
    try:
        problem.start = dict(problem.start, t=saved + d)
        j1 = problem.forward(c)
    finally:
        problem.start = dict(problem.start, t=saved)
    assert abs((j1 - j0) - float(np.sum(gt0 * d))) <= 1e-8 * abs(j1 - j0)


def test_checkpoint_spacing_and_final_state(problem):
    c = _controls(8, seed=5)
    problem.forward(c)
    t_end = problem.cav.get("t")
    g20 = problem.gradient()
    assert np.array_equal(problem.cav.get("t"), t_end)
    problem.every = 7  # 60 = 8 * 7 + 4
    try:
        problem.forward(c)
        g7 = problem.gradient()
    finally:
        problem.every = 20
    np.testing.assert_allclose(g7, g20, rtol=1e-12, atol=1e-15)


def test_autograd_function(problem):
    torch = pytest.importorskip("torch")
    c = torch.tensor(_controls(8, seed=6), dtype=torch.float64, requires_grad=True)
    nu = A.HotWallNusselt.apply(c, problem)
    nu.backward()
    ref = problem.gradient()
    np.testing.assert_allclose(c.grad.numpy(), ref, rtol=1e-14, atol=0)


# ---- adjoint of the whole step (ixh_adjoint_coupled)

def test_coupled_equals_energy_adjoint_without_buoyancy(problem):
    """Without buoyancy the velocity adjoints never return to T: the coupled
    gradient equals the energy-only one."""
    c = _controls(8, seed=7)
    problem.forward(c)
    problem.coupled = False
    g_e = problem.gradient()
    problem.coupled = True
    g_c = problem.gradient()
    np.testing.assert_allclose(g_c, g_e, rtol=1e-11, atol=1e-14 * np.max(np.abs(g_e)))


@pytest.fixture(scope="module")
def buoyant():
    from ixchel2d_amd.simple import Cavity
    case = C.natural_convection_case(24, 30, height=1.3, Ra=3e4, iters=40, dt=0.01, seed=3, beta=1.1)
    warm = Cavity(case, solver="pcr")
    warm.run(150)
    start = {f: warm.get(f) for f in ("u", "v", "t", "p")}
    return A.HotWallProblem(case, ncontrols=6, steps=40, every=15, start=start, coupled=True)


def test_coupled_adjoint_against_central_differences_with_buoyancy(buoyant):
    c = _controls(6, seed=8)
    buoyant.forward(c)
    g = buoyant.gradient()
    buoyant.coupled = False
    g_frozen = buoyant.gradient()
    buoyant.coupled = True
    h = 1e-6
    fd = np.zeros_like(c)
    for i in range(len(c)):
        e = np.zeros_like(c)
        e[i] = h
        fd[i] = (buoyant.forward(c + e) - buoyant.forward(c - e)) / (2 * h)
    err = np.max(np.abs(g - fd)) / np.max(np.abs(fd))
    err_frozen = np.max(np.abs(g_frozen - fd)) / np.max(np.abs(fd))
    assert err <= 1e-5, (err, err_frozen, g, fd)
    assert err < err_frozen


def test_coupled_checkpoint_spacing_with_buoyancy(buoyant):
    """With buoyancy the velocity adjoints feed back into T, so every entry
    of the recomputed tape (u, v at each phase of each step) reaches the
    gradient; another checkpoint spacing (40 = 5 * 7 + 5 steps) moves the
    segment starts, where the state comes from a checkpoint, and must leave
    the gradient unchanged."""
    c = _controls(6, seed=11)
    buoyant.forward(c)
    g15 = buoyant.gradient()
    buoyant.every = 7
    try:
        buoyant.forward(c)
        g7 = buoyant.gradient()
    finally:
        buoyant.every = 15
    np.testing.assert_allclose(g7, g15, rtol=1e-12, atol=1e-15)


def test_coupled_initial_temperature_with_buoyancy(buoyant):
    c = _controls(6, seed=9)
    buoyant.forward(c)
    _, gt0 = buoyant.cav.adjoint_energy(buoyant.W, coupled=True)
    rng = np.random.default_rng(10)
    d = rng.standard_normal(buoyant.start["t"].shape)
    eps = 1e-5
    saved = buoyant.start["t"]
    try:
        buoyant.start = dict(buoyant.start, t=saved + eps * d)
        jp = buoyant.forward(c)
        buoyant.start = dict(buoyant.start, t=saved - eps * d)
        jm = buoyant.forward(c)
    finally:
        buoyant.start = dict(buoyant.start, t=saved)
    fd = (jp - jm) / (2 * eps)
    assert abs(fd - float(np.sum(gt0 * d))) <= 1e-5 * abs(fd)


def test_autograd_backward_after_a_second_forward(buoyant):
    """Two evaluations, then the backward pass of the first: its gradient is
    that of its own controls, not of the run the cavity recorded last (with
    buoyancy, where the gradient depends on the controls)."""
    torch = pytest.importorskip("torch")
    c1 = torch.tensor(_controls(6, seed=12), dtype=torch.float64, requires_grad=True)
    c2 = torch.tensor(_controls(6, seed=13), dtype=torch.float64, requires_grad=True)
    nu1 = A.HotWallNusselt.apply(c1, buoyant)
    nu2 = A.HotWallNusselt.apply(c2, buoyant)
    nu1.backward()
    buoyant.forward(c1.detach().numpy())
    g1 = buoyant.gradient()
    np.testing.assert_allclose(c1.grad.numpy(), g1, rtol=1e-14, atol=0)
    nu2.backward()
    buoyant.forward(c2.detach().numpy())
    g2 = buoyant.gradient()
    np.testing.assert_allclose(c2.grad.numpy(), g2, rtol=1e-14, atol=0)
    assert not np.allclose(g1, g2, rtol=1e-6, atol=0)


# ---- the adjoint workload at its own size: natural convection on 129^2
# cells (uniform mesh), Ra = 1e5, Pr = 0.71, 64 piecewise-linear hot-wall
# controls 1 +- U(0.1) (seed 0), J = mean hot-wall Nu after K = 400
# iterations from a converged start, checkpoints every 20

@pytest.fixture(scope="module")
def nc129():
    from ixchel2d_amd.simple import Cavity
    case = C.natural_convection_case(130, 130, height=1.0, Ra=1e5, iters=400, dt=0.01, seed=0,
                                     mesh_kind="unif")
    warm = Cavity(case, solver="pcr")
    warm.run(20000, cap=1)
    h = warm.run(1)
    start = {f: warm.get(f) for f in ("u", "v", "t", "p")}
    warm.close()
    assert h["residuo"][-1] < 1e-5, h["residuo"][-1]
    prob = A.HotWallProblem(case, ncontrols=64, steps=400, every=20, start=start, coupled=True)
    c = 1.0 + np.random.default_rng(0).uniform(-0.1, 0.1, 64)
    prob.forward(c)
    return prob, c, prob.gradient()


FD_CONTROLS = (0, 13, 31, 50, 63)


def test_adjoint129_against_gpu_central_differences(nc129):
    prob, c, g = nc129
    h = 1e-6
    fd = []
    for i in FD_CONTROLS:
        e = np.zeros(64)
        e[i] = h
        fd.append((prob.forward(c + e) - prob.forward(c - e)) / (2 * h))
    fd = np.array(fd)
    err = np.max(np.abs(g[list(FD_CONTROLS)] - fd)) / np.max(np.abs(fd))
    assert err <= 1e-5, (err, g[list(FD_CONTROLS)], fd)


def test_adjoint129_against_oracle_central_difference(nc129):
    """The FP64 reference: central differences of the oracle's own K-step
    run (the C restatement on the CPU) for one control."""
    prob, c, g = nc129
    i = 31
    h = 1e-6

    def J(cc):
        o = Oracle(prob.case)
        o.set_bc("t", "a", prob.types_a, prob.wall_values(cc))
        for f, arr in prob.start.items():
            o.set(f, arr)
            if f in ("u", "v", "t"):
                o.set(f + "_ant", arr)
        o.run(400)
        return o.nusselt_x()[0]

    e = np.zeros(64)
    e[i] = h
    ref = (J(c + e) - J(c - e)) / (2 * h)
    assert abs(g[i] - ref) <= 1e-5 * abs(ref), (g[i], ref)
