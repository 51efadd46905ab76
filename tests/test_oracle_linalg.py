"""Oracle band LU (LAPACK gbtrf/gbtrs restatement) against numpy / scipy."""
import numpy as np
import pytest
import scipy.linalg as sl


def _band(n, kl, ku, rng, diag=0.0):
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - kl), min(n, i + ku + 1)):
            A[i, j] = rng.normal()
        A[i, i] += diag
    return A


@pytest.mark.parametrize("n,kl,ku", [(1, 0, 0), (7, 2, 3), (40, 5, 5), (200, 17, 17), (64, 0, 4), (64, 4, 0)])
def test_gbtrf_matches_dense(oracle_lib, n, kl, ku):
    rng = np.random.default_rng(n * 100 + kl)
    A = _band(n, kl, ku, rng)  # not diagonally dominant: row interchanges happen
    b = rng.normal(size=n)
    info, x = oracle_lib.gbsolve(A, kl, ku, b)
    assert info == 0
    xr = np.linalg.solve(A, b)
    assert np.allclose(x, xr, rtol=1e-9, atol=1e-9 * np.abs(xr).max())


def test_gbtrf_matches_scipy_solve_banded(oracle_lib):
    rng = np.random.default_rng(7)
    n, kl, ku = 300, 12, 9
    A = _band(n, kl, ku, rng, diag=0.5)
    ab = np.zeros((kl + ku + 1, n))
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            ab[ku + i - j, j] = A[i, j]
    b = rng.normal(size=n)
    info, x = oracle_lib.gbsolve(A, kl, ku, b)
    assert info == 0
    xs = sl.solve_banded((kl, ku), ab, b)
    assert np.allclose(x, xs, rtol=1e-10, atol=1e-10 * np.abs(xs).max())


def test_gbtrf_needs_pivoting(oracle_lib):
    # zero leading pivot: only solvable with a row interchange
    A = np.array([[0.0, 1.0, 0.0], [2.0, 3.0, 1.0], [0.0, 1.0, 4.0]])
    b = np.array([1.0, 2.0, 3.0])
    info, x = oracle_lib.gbsolve(A, 1, 1, b)
    assert info == 0
    assert np.allclose(A @ x, b)


def test_gbtrf_reports_singular(oracle_lib):
    A = np.array([[1.0, 2.0], [2.0, 4.0]])
    info, _ = oracle_lib.gbsolve(A, 1, 1, np.ones(2))
    assert info != 0
