"""CPU checks of the mixed-size batch generator: deterministic from its
seed, sizes from the four shapes, half lid-driven and half natural
convection, cylinders in 2 of every 16 cavities with the body inside the
cavity and its temperature pinned by the penalty."""
import numpy as np

from ixchel2d_amd import case as C


def test_batch_cases_are_deterministic_and_mixed():
    a = C.batch_cases(32, seed=3)
    b = C.batch_cases(32, seed=3)
    assert [c.name for c in a] == [c.name for c in b]
    shapes = {(c.mesh.mi, c.mesh.nj) for c in a}
    assert shapes <= {(m, n) for m, n, _ in C.BATCH_SIZES}
    assert sum(c.params.adimen == "mixta" for c in a) == 16
    cyl = [k for k, c in enumerate(a) if c.sources]
    assert cyl == [3, 11, 19, 27]
    for k in cyl:
        c = a[k]
        sc, sp = c.sources["t"]
        inside = sp != 0.0
        assert inside.any()
        np.testing.assert_allclose(sc[inside] / -sp[inside], 1.0)
        # the body stays clear of the walls: ghost rows and columns untouched
        assert not inside[0, :].any() and not inside[-1, :].any()
        assert not inside[:, 0].any() and not inside[:, -1].any()
