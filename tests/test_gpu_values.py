"""Device-resident value updates (pastix_amd_values_dev): factoring new values
gathered on the device into the permuted CSC gives the factors of a fresh
host-side numfact with those values, for a symmetric matrix stored by its
lower triangle or by both triangles and for a nonsymmetric one; ||A||_F is
recomputed on the device; the solve after the update solves the new system."""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.sparse as sp

import pastix_amd as pa
from pastix_amd import problems, _lib

pytestmark = pytest.mark.gpu


def _handle(ft):
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = ft
    ip[pa.iparm.verbose] = pa.verbose.Not
    return pa.init(ip, dp), ip, dp


def _pair(kind, grid):
    """Two matrices with the same stored pattern and different values."""
    if kind == "lu":
        A1 = problems.convection_diffusion_27(*grid)
        A2 = problems.convection_diffusion_27(*grid, velocity=(0.9, 0.4, 0.25), diffusion=1.7)
        ft, mt = pa.factotype.LU, pa.mtxtype.General
    else:
        th1 = problems.anisotropic_conductivity(*grid, seed=3)
        th2 = problems.anisotropic_conductivity(*grid, seed=4)
        A1 = problems.anisotropic_diffusion(*grid, th1)
        A2 = problems.anisotropic_diffusion(*grid, th2)
        if kind == "llt-lower":
            A1, A2 = sp.tril(A1), sp.tril(A2)
        ft, mt = pa.factotype.LLT, pa.mtxtype.Symmetric
    A1, A2 = sp.csc_matrix(A1), sp.csc_matrix(A2)
    A1.sort_indices()
    A2.sort_indices()
    assert np.array_equal(A1.indices, A2.indices) and np.array_equal(A1.indptr, A2.indptr)
    return A1, A2, ft, mt


def _frobenius(A, lower_only):
    """||A||_F of the full matrix, summed exactly (math.fsum)."""
    C = sp.coo_matrix(A)
    w = np.where((C.row != C.col) & lower_only, 2.0, 1.0)
    return math.sqrt(math.fsum(w * C.data * C.data))


@pytest.mark.parametrize("kind", ["llt-lower", "llt-full", "lu"])
def test_device_values_match_host_numfact(kind):
    import torch
    grid = (11, 9, 8)
    A1, A2, ft, mt = _pair(kind, grid)
    lib = _lib.load()
    h1, ip1, dp1 = _handle(ft)
    h2, ip2, dp2 = _handle(ft)
    try:
        s1 = pa.spmatrix(A1, mtxtype=mt)
        s2 = pa.spmatrix(A2, mtxtype=mt)
        pa.task_analyze(h1, s1, pa.order.grid(*grid))
        pa.task_analyze(h2, s2, pa.order.grid(*grid))
        pa.task_numfact(h1, s1)
        pa.task_numfact(h2, s2)                     # reference: host numfact of A2
        norm_ref = _frobenius(A2, kind == "llt-lower")
        assert abs(float(dp2[pa.dparm.a_norm]) - norm_ref) <= 1e-12 * norm_ref
        for rep in range(2):                        # the second call reuses the value map
            vals = A2 if rep == 0 else A1
            d_v = torch.from_numpy(np.ascontiguousarray(vals.data)).cuda()
            torch.cuda.synchronize()
            assert lib.pastix_amd_values_dev(h1, C.c_void_p(d_v.data_ptr())) == 0
            assert lib.pastix_subtask_bcsc2ctab(h1) == 0
            assert lib.pastix_subtask_sopalin(h1) == 0
        # back to A2 and compare the factors with the host numfact of A2
        d_v = torch.from_numpy(np.ascontiguousarray(A2.data)).cuda()
        torch.cuda.synchronize()
        assert lib.pastix_amd_values_dev(h1, C.c_void_p(d_v.data_ptr())) == 0
        assert abs(float(dp1[pa.dparm.a_norm]) - norm_ref) <= 1e-12 * norm_ref
        assert lib.pastix_subtask_bcsc2ctab(h1) == 0
        assert lib.pastix_subtask_sopalin(h1) == 0
        for side in ((0, 1) if ft == pa.factotype.LU else (0,)):
            L1, L2 = pa.get_coeftab(h1, side), pa.get_coeftab(h2, side)
            assert np.abs(L1 - L2).max() <= 1e-13 * np.abs(L2).max(), side
        # the solve after the update solves A2 x = b
        Afull = A2 if kind != "llt-lower" else sp.csc_matrix(A2 + sp.tril(A2, -1).T)
        n = Afull.shape[0]
        b = problems.uniform(n, seed=9)
        x = b.copy()
        pa.task_solve(h1, s1, x)
        assert problems.backward_error(Afull, x, b) <= 1e-14
    finally:
        pa.finalize(h1)
        pa.finalize(h2)


def test_device_values_before_numfact_is_an_error():
    import torch
    h, ip, dp = _handle(pa.factotype.LLT)
    try:
        grid = (4, 4, 4)
        A = problems.laplacian(*grid)
        s = pa.spmatrix(A, mtxtype=pa.mtxtype.Symmetric)
        pa.task_analyze(h, s, pa.order.grid(*grid))
        d_v = torch.zeros(A.nnz, dtype=torch.float64, device="cuda")
        assert _lib.load().pastix_amd_values_dev(h, C.c_void_p(d_v.data_ptr())) != 0
    finally:
        pa.finalize(h)
