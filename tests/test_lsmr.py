"""CPU checks of the LSMR oracle (src/shared/lsmrModule_csr.f90 with the
products of src/shared/parallel.cpp): the plain-C restatement gives the same
bits as the compiled reference run serially, and both land on the solution
the mathematics fixes -- the least-squares solution of an overdetermined
system, the damped normal equations (A^T A + damp^2 I) x = A^T b, and the
minimum-norm solution of an underdetermined one -- to the accuracy LSMR's
stop rules (atol = btol = 1e-5) allow in float."""
import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.skipif(not pyoracle.available("rest"), reason="oracle not built")


def sparse_problem(m, n, density, seed):
    """A random sparse system with a strengthened diagonal (well conditioned),
    as dense float32 and as CSR arrays (0-based, row order)."""
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(m, n)) * (rng.random((m, n)) < density)
    k = min(m, n)
    A[np.arange(k), np.arange(k)] += 3.0
    A = A.astype(np.float32)
    ptr = np.concatenate([[0], np.cumsum((A != 0).sum(1))]).astype(np.int32)
    ind = np.concatenate([np.nonzero(r)[0] for r in A]).astype(np.int32)
    val = np.concatenate([r[r != 0] for r in A]).astype(np.float32)
    b = rng.normal(size=m).astype(np.float32)
    return A, ptr, ind, val, b


CASES = [(300, 120, 0.05, 0.0), (300, 120, 0.05, 0.5), (80, 200, 0.1, 0.0), (2000, 700, 0.01, 0.1)]


@pytest.mark.parametrize("m,n,density,damp", CASES)
def test_restatement_matches_solution(m, n, density, damp):
    A, ptr, ind, val, b = sparse_problem(m, n, density, m + n)
    x, info = pyoracle.lsmr(m, n, val, ind, ptr, b, damp=damp)
    Ad = A.astype(np.float64)
    if damp > 0:
        want = np.linalg.solve(Ad.T @ Ad + damp ** 2 * np.eye(n), Ad.T @ b)
    else:
        want = np.linalg.lstsq(Ad, b.astype(np.float64), rcond=None)[0]  # min-norm if m < n
    assert info["istop"] in (1, 2, 3)
    assert np.abs(x - want).max() <= 1e-3 * np.abs(want).max()
    # the returned norm estimates describe that solution
    assert abs(info["normx"] - np.linalg.norm(x)) <= 1e-4 * np.linalg.norm(x)


@pytest.mark.parametrize("m,n,density,damp", CASES)
def test_restatement_matches_reference_bits(m, n, density, damp):
    if not __import__("os").path.exists(pyoracle.LIB_LSMR_REF):
        pytest.skip("reference LSMR not built")
    A, ptr, ind, val, b = sparse_problem(m, n, density, m + n)
    x, info = pyoracle.lsmr(m, n, val, ind, ptr, b, damp=damp, kind="rest")
    xr, ir = pyoracle.lsmr(m, n, val, ind, ptr, b, damp=damp, kind="ref", threads=1)
    assert np.array_equal(x, xr)
    assert info == ir


def test_iteration_limit_and_zero_rhs():
    A, ptr, ind, val, b = sparse_problem(200, 80, 0.05, 7)
    x, info = pyoracle.lsmr(200, 80, val, ind, ptr, b, itnlim=3)
    assert info["istop"] == 7 and info["itn"] == 3
    x0, info0 = pyoracle.lsmr(200, 80, val, ind, ptr, np.zeros(200, np.float32))
    assert not x0.any() and info0["itn"] == 0


def test_reference_runs_at_frechet_width():
    """The compiled reference keeps its work vectors on the stack
    (lsmrModule_csr.f90:789); at a million model cells the wrapper's thread
    stack must hold them.  The restatement agrees there too."""
    if not __import__("os").path.exists(pyoracle.LIB_LSMR_REF):
        pytest.skip("reference LSMR not built")
    m, n = 3000, 1_032_256
    rng = np.random.default_rng(11)
    ptr = np.arange(0, 2 * m + 1, 2, dtype=np.int32)
    ind = rng.integers(0, n, 2 * m).astype(np.int32)
    val = rng.normal(size=2 * m).astype(np.float32)
    b = rng.normal(size=m).astype(np.float32)
    xr, ir = pyoracle.lsmr(m, n, val, ind, ptr, b, itnlim=4, kind="ref", threads=2)
    x, info = pyoracle.lsmr(m, n, val, ind, ptr, b, itnlim=4, kind="rest")
    assert ir["itn"] == info["itn"]
    assert np.abs(x - xr).max() <= 1e-5 * np.abs(xr).max()


@pytest.mark.parametrize("m,n,density,damp", CASES[:3])
def test_double_twin_reaches_the_exact_solution(m, n, density, damp):
    """The double twin (a yardstick for float iterates) converges to the
    least-squares / damped / minimum-norm solution to near double rounding."""
    A, ptr, ind, val, b = sparse_problem(m, n, density, m + n)
    x, info = pyoracle.lsmr(m, n, val, ind, ptr, b, damp=damp, atol=1e-14, btol=1e-14,
                            conlim=1e14, kind="f64")
    Ad = A.astype(np.float64)
    if damp > 0:
        want = np.linalg.solve(Ad.T @ Ad + damp ** 2 * np.eye(n), Ad.T @ b)
    else:
        want = np.linalg.lstsq(Ad, b.astype(np.float64), rcond=None)[0]
    assert x.dtype == np.float64
    assert np.abs(x - want).max() <= 1e-10 * np.abs(want).max()
