"""The block-tridiagonal cyclic-reduction solver (gf_btdsystem) against sparse
direct solves (scipy), including blocks that need row interchanges (the branch-free
sweep hands them to the pivoting one) and a singular system."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from giraffe_amd import runtime

pytestmark = pytest.mark.gpu


def _sparse(A, B, C):
    nblk, m = B.shape[0], B.shape[1]
    blocks = [[None] * nblk for _ in range(nblk)]
    for i in range(nblk):
        blocks[i][i] = B[i]
        if i > 0:
            blocks[i][i - 1] = A[i]
        if i + 1 < nblk:
            blocks[i][i + 1] = C[i]
    return sp.bmat(blocks, format="csc")


def _system(m, nblk, rng, permuted=False):
    A = rng.uniform(-1, 1, (nblk, m, m))
    C = rng.uniform(-1, 1, (nblk, m, m))
    B = rng.uniform(-1, 1, (nblk, m, m)) + 4.0 * m * np.eye(m)
    if permuted:
        # row-permuted diagonal blocks: small diagonal entries, large off-diagonal
        # ones, so the threshold rule takes row interchanges
        for i in range(nblk):
            B[i] = B[i][rng.permutation(m)]
    return A, B, C


@pytest.mark.parametrize("m,nblk", [(12, 300), (16, 1000), (20, 97), (6, 5000)])
@pytest.mark.parametrize("permuted", [False, True])
def test_btdsystem_matches_dense(m, nblk, permuted):
    rng = np.random.default_rng(m * 1000 + nblk + permuted)
    A, B, C = _system(m, nblk, rng, permuted)
    b = rng.uniform(-1, 1, nblk * m)
    x, fail = runtime.btdsystem(A, B, C, b)
    assert fail == 0
    K = _sparse(A, B, C)
    ref = spla.spsolve(K, b)
    assert np.abs(x - ref).max() <= 1e-10 * np.abs(ref).max()
    assert np.abs(K @ x - b).max() <= 1e-10 * np.abs(b).max() * abs(K).max()


def test_btdsystem_reports_singular_block():
    rng = np.random.default_rng(7)
    A, B, C = _system(16, 200, rng)
    A[:] = 0.0
    C[:] = 0.0
    B[37] = 0.0  # decoupled zero block: no pivot can be found
    x, fail = runtime.btdsystem(A, B, C, rng.uniform(-1, 1, 200 * 16))
    assert fail == 1
