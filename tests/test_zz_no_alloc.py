"""ixh_tridiagonal_noalloc (the entry point that allocates no device memory,
working through a static device scratch array) against ixh_tridiagonal,
through the C-ABI: the same solutions and overwritten pivots bit for bit,
for both solvers, on two calls in a row (the second smaller, so stale
scratch contents from the first would show), and the size limit rejected
on the host.  Runs last: a failure here hides no other test."""
import ctypes as C

import numpy as np
import pytest

from ixchel2d_amd import _lib

pytestmark = pytest.mark.gpu

LIMIT = 1 << 18  # values per scratch array


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)


def _systems(n, nsys, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1.0, 0.0, (nsys, n))
    c = rng.uniform(-1.0, 0.0, (nsys, n))
    b = 2.5 + rng.uniform(0.0, 1.0, (nsys, n))
    r = rng.standard_normal((nsys, n))
    return a, b, c, r


def _call(fn, a, b, c, r, solver):
    dp = C.POINTER(C.c_double)
    a = np.ascontiguousarray(a)
    c = np.ascontiguousarray(c)
    b = np.ascontiguousarray(b).copy()
    r = np.ascontiguousarray(r).copy()
    nsys, n = r.shape
    rc = fn(a.ctypes.data_as(dp), b.ctypes.data_as(dp), c.ctypes.data_as(dp), r.ctypes.data_as(dp),
            n, nsys, _lib.SOLVER_ID[solver])
    return rc, r, b


@pytest.mark.parametrize("solver", ["pcr", "thomas"])
def test_noalloc_equals_allocating_entry_point(solver):
    L = _lib.load()
    # first call near the scratch limit, then a smaller one
    for n, nsys, seed in ((700, LIMIT // 700, 1), (301, 37, 2)):
        a, b, c, r = _systems(n, nsys, seed)
        rc0, x0, b0 = _call(L.ixh_tridiagonal, a, b, c, r, solver)
        rc1, x1, b1 = _call(L.ixh_tridiagonal_noalloc, a, b, c, r, solver)
        assert rc0 == 0 and rc1 == 0, L.ixh_last_error()
        assert np.array_equal(x1, x0), f"n={n} nsys={nsys}: solutions differ"
        assert np.array_equal(b1, b0), f"n={n} nsys={nsys}: pivots differ"
        # and it is a solution (dense check on a few systems)
        for k in (0, nsys - 1):
            A = np.diag(b[k]) + np.diag(a[k, 1:], -1) + np.diag(c[k, :-1], 1)
            np.testing.assert_allclose(A @ x1[k], r[k], rtol=1e-12, atol=1e-12)


def test_noalloc_rejects_sizes_beyond_its_scratch():
    L = _lib.load()
    n = 1024
    nsys = LIMIT // n + 1
    a, b, c, r = _systems(n, nsys, 3)
    rc, _, _ = _call(L.ixh_tridiagonal_noalloc, a, b, c, r, "pcr")
    assert rc == -4
    assert b"scratch" in L.ixh_last_error()
