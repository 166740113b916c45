"""CPU checks of the adjoint harness's host-side pieces: the Nusselt
weights reproduce the oracle's vertical-wall Nusselt number (a linear
functional of T), and the control matrix interpolates linearly."""
import numpy as np

from ixchel2d_amd import adjoint as A
from ixchel2d_amd import case as C
from oracle.pyoracle import Oracle


def test_nusselt_weights_match_oracle():
    case = C.natural_convection_case(20, 30, height=1.5, Ra=1e4, iters=5, dt=0.01, seed=2)
    o = Oracle(case)
    o.run(5)
    t = o.get("t")
    W = A.nusselt_x_weights(case.mesh)
    nu_w, _ = o.nusselt_x()
    assert abs(float(np.sum(W * t)) - nu_w) <= 1e-12 * max(1.0, abs(nu_w))
    We = A.nusselt_x_weights(case.mesh, wall="east")
    assert abs(float(np.sum(We * t)) - o.nusselt_x()[1]) <= 1e-12 * max(1.0, abs(o.nusselt_x()[1]))


def test_control_matrix_is_linear_interpolation():
    m = C.make_mesh(16, 40, 1.0, 2.0, "tanh", 1.2)
    B = A.control_matrix(m, 5)
    np.testing.assert_allclose(B.sum(axis=1), 1.0, rtol=0, atol=1e-15)
    yc = np.linspace(m.yv[0], m.yv[-1], 5)
    # a linear profile in y is reproduced exactly between the end controls
    g = B @ (3.0 + 2.0 * yc)
    inside = (m.yp >= yc[0]) & (m.yp <= yc[-1])
    np.testing.assert_allclose(g[inside], 3.0 + 2.0 * m.yp[inside], rtol=1e-14)
