"""Batched pass over many uneven SPD systems (pastix_amd.batch): one
analysis of the block-diagonal forest, one factorization launch sequence and
one multi-RHS solve; each system's solution is checked against an
independent sparse direct solve, and the small ones against dense LAPACK."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from pastix_amd import batch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mixed", [False], ids=["fp64"])
def test_batched_uneven_laplacians(mixed):
    mats, grids = batch.uneven_laplacians(count=48, low=8, high=72, seed=1)
    rng = np.random.default_rng(1)
    rhs = [rng.uniform(-1, 1, (m.shape[0], 16)) for m in mats]
    s = batch.BatchedSolver(mats, grids, mixed=mixed)
    try:
        s.factorize()
        xs = s.solve(rhs)
    finally:
        s.close()
    for A, b, x in zip(mats, rhs, xs):
        ref = spla.spsolve(A.tocsc(), b)
        assert np.abs(x - ref).max() <= 1e-12 * np.abs(ref).max()
        if A.shape[0] <= 4096:
            dense = np.linalg.solve(A.toarray(), b)
            assert np.abs(x - dense).max() <= 1e-12 * np.abs(dense).max()
