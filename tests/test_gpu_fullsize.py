"""Parity at the full sizes of the measured workloads: the FP64 factors of the
100^3 Laplacian (LL^t) and of the 80^3 convection-diffusion matrix (LU) equal
the CPU oracle's on the same structure, and the solution of the 100^3 system
equals the oracle's solve; the batched pass of 256 uneven systems solves every
system to backward error 1e-14.

The oracle runs its level-parallel driver with per-task BLAS on the host cores
(pinned to the sequential oracle by test_cpu_oracle.py), so each case takes
seconds on the CPU."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import pastix_amd as pa
from pastix_amd import problems
from oracle import pyoracle as orc
from helpers import Problem

pytestmark = pytest.mark.gpu


def _threads():
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return 4


def _masked_relerr(G, O, mask):
    """max |G - O| over the factor entries / max |O|, in place on G."""
    scale = np.abs(O).max()
    np.subtract(G, O, out=G)
    G[~mask] = 0.
    return np.abs(G).max() / scale


def test_laplacian_100_fp64_factors_and_solution_match_the_oracle():
    grid = (100, 100, 100)
    print("case laplacian 100^3 LL^t FP64", flush=True)
    A = problems.laplacian(*grid)
    P = Problem(A, grid, pa.factotype.LLT)
    try:
        n = P.n
        xt = problems.uniform(n, seed=0)
        b = A @ xt
        x = pa.task_solve(P.h, P.spm, b.copy())
        Lg = pa.get_coeftab(P.h, 0)
        Lo, _, npiv, sec = orc.factorize_parallel(P.S, P.A, P.perm, P.crit, _threads(), use_blas=True)
        assert npiv == P.iparm[pa.iparm.static_pivoting] == 0
        xo_p = orc.solve(P.S, 0, Lo, None, b[P.invp])
        xo = np.empty(n)
        xo[P.invp] = xo_p
        assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max()
        assert problems.backward_error(A, x, b) <= 1e-14
        assert _masked_relerr(Lg, Lo, P.lower_mask()) <= 1e-12
    finally:
        P.close()


def test_convection_diffusion_80_fp64_lu_factors_match_the_oracle():
    grid = (80, 80, 80)
    print("case convection-diffusion 80^3 LU FP64", flush=True)
    A = problems.convection_diffusion_27(*grid)
    P = Problem(A, grid, pa.factotype.LU)
    try:
        Lg, Ug = P.gpu_factors()
        Lo, Uo, npiv, sec = orc.factorize_parallel(P.S, P.A, P.perm, P.crit, _threads(), use_blas=True,
                                                   factotype=2)
        assert npiv == P.iparm[pa.iparm.static_pivoting]
        allm = np.ones(Lg.shape[0], dtype=bool)
        assert _masked_relerr(Lg, Lo, allm) <= 1e-12
        assert _masked_relerr(Ug, Uo, allm) <= 1e-12
    finally:
        P.close()


def test_batch_of_256_uneven_systems_full_size():
    from pastix_amd import batch
    print("case batch 256 systems", flush=True)
    mats, grids = batch.uneven_laplacians(count=256, low=64, high=512, seed=1)
    rng = np.random.default_rng(1)
    rhs = [rng.uniform(-1, 1, (m.shape[0], 16)) for m in mats]
    s = batch.BatchedSolver(mats, grids)
    try:
        s.factorize()
        xs = s.solve(rhs)
    finally:
        s.close()
    worst = 0.0
    for A, b, x in zip(mats, rhs, xs):
        r = b - A @ x
        nA = abs(A).sum(axis=1).max()
        worst = max(worst, float(np.abs(r).max() / (nA * np.abs(x).max() + np.abs(b).max())))
    assert worst <= 1e-14, worst
    for i in (0, 100, 255):
        ref = spla.spsolve(sp.csc_matrix(mats[i]), rhs[i])
        assert np.abs(xs[i] - ref).max() <= 1e-12 * np.abs(ref).max()
