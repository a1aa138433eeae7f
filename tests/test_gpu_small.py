"""The fused one-wave kernel of the narrow cblks (k_small_llt) gives the same
factors as the general diagonal / TRSM / scatter-GEMM kernels (switched off
with PASTIX_AMD_SMALL=0), in FP64 and FP32, on 2D and 3D grids where most
cblks are narrow; both agree with the CPU oracle, and so do the multi-RHS
and one-RHS solves of the narrow cblks (k_small_solve, two cblks of at most
32 rows per wavefront) and of the general kernels."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import pastix_amd as pa
from pastix_amd import problems
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def _factor(A, grid, mixed, small):
    old = os.environ.get("PASTIX_AMD_SMALL")
    os.environ["PASTIX_AMD_SMALL"] = "1" if small else "0"
    try:
        ip, dp = pa.initParam()
        ip[pa.iparm.factorization] = pa.factotype.LLT
        ip[pa.iparm.verbose] = pa.verbose.Not
        if mixed:
            ip[pa.iparm.mixed] = 1
        h = pa.init(ip, dp)
        spm = pa.spmatrix(A, mtxtype=pa.mtxtype.Symmetric)
        pa.task_analyze(h, spm, pa.order.grid(*grid))
        pa.task_numfact(h, spm)
    finally:
        if old is None:
            del os.environ["PASTIX_AMD_SMALL"]
        else:
            os.environ["PASTIX_AMD_SMALL"] = old
    return h, spm, ip


def _lower_mask(cblk, size):
    keep = np.ones(size, dtype=bool)
    for f, l, fb, st, off in cblk:
        nc = l - f + 1
        t = keep[off:off + nc * nc].reshape((nc, nc))
        t[np.tril_indices(nc, -1)] = False   # (col, row) view: drop the upper triangle
    return keep


@pytest.mark.parametrize("grid,mixed", [((60, 50, 1), False), ((60, 50, 1), True),
                                        ((13, 12, 11), False), ((13, 12, 11), True)],
                         ids=["2d-fp64", "2d-fp32", "3d-fp64", "3d-fp32"])
def test_narrow_cblk_kernel_matches_general_kernels(grid, mixed):
    A = problems.laplacian(*grid, shift=0.1)
    h1, s1, ip1 = _factor(A, grid, mixed, True)
    h0, s0, ip0 = _factor(A, grid, mixed, False)
    try:
        cblk, blok, perm, invp, dims = pa.get_structure(h1)
        nc = cblk[:, 1] - cblk[:, 0] + 1
        assert (nc <= 16).mean() > 0.5          # the narrow path is exercised
        # narrow cblks of at most 32 rows: solved two per wavefront
        assert ((nc <= 16) & (cblk[:, 3] <= 32)).sum() > 0
        keep = _lower_mask(cblk, int(dims[3]))
        L1, L0 = pa.get_coeftab(h1, 0)[keep], pa.get_coeftab(h0, 0)[keep]
        tol = 2e-6 if mixed else 1e-13
        assert np.abs(L1 - L0).max() <= tol * np.abs(L0).max()
        assert ip1[pa.iparm.static_pivoting] == ip0[pa.iparm.static_pivoting]
        if not mixed:
            S = orc.Structure(cblk, blok)
            crit = np.finfo(np.float64).eps / 2 * sp.linalg.norm(A)
            Lo, _, _ = orc.factorize(S, A, perm, 0, crit)
            assert np.abs(L1 - Lo[keep]).max() <= 1e-12 * np.abs(Lo).max()
        n = A.shape[0]
        B = np.asfortranarray(np.random.default_rng(2).uniform(-1, 1, (n, 5)))
        X1, X0 = B.copy(order="F"), B.copy(order="F")
        pa.task_solve(h1, s1, X1)                  # narrow cblks: k_small_solve, 5 RHS at once
        pa.task_solve(h0, s0, X0)
        assert np.abs(X1 - X0).max() <= (1e-4 if mixed else 1e-12) * np.abs(X0).max()
        for j in range(5):
            assert problems.backward_error(A, X1[:, j], B[:, j]) <= (1e-6 if mixed else 1e-14)
        x1 = B[:, :1].copy(order="F")
        pa.task_solve(h1, s1, x1)                  # one RHS: the two-array sweeps
        assert np.abs(x1[:, 0] - X0[:, 0]).max() <= (1e-4 if mixed else 1e-12) * np.abs(X0).max()
    finally:
        pa.finalize(h1)
        pa.finalize(h0)
