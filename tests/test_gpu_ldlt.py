"""LDL^t (IPARM_FACTORIZATION = PastixFactLDLT / LDLH) on the GPU against the
oracle's restatement of sequential_zsytrf (oracle_*sytrf, pinned to the
unique no-pivot LDL^t in test_cpu_oracle.py): the L array (L, D and D L^t
in the diagonal blocks), the L D copy of the panels, the L / D / L^t solve,
the solve subtasks with pastix_subtask_diag, and FP32 factors refined in
FP64."""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import pastix_amd as pa
from pastix_amd import _lib, problems
from oracle import pyoracle as orc
from helpers import Problem

pytestmark = pytest.mark.gpu

LEFT = 141
LOWER = 122
NONUNIT, UNIT = 131, 132
NOTRANS, TRANS = 111, 112


def _indefinite(grid, shift):
    """The Laplacian shifted into indefiniteness: some Schur complements of
    the elimination come close to singular, so pivots are small and the
    factors grow (max |L| ~ 300 times the matrix entries at shift 2.5):
    rounding differences are amplified accordingly."""
    n = int(np.prod(grid))
    return (problems.laplacian(*grid) - shift * sp.eye(n)).tocsc()


def _indefinite_dd(grid, seed):
    """Symmetric, indefinite and strictly diagonally dominant: the 7-point
    stencil's -1 couplings with a diagonal of +7 or -7 per cell (random
    signs).  Every Schur complement of a no-pivot elimination stays strictly
    diagonally dominant, so the pivots keep a magnitude of at least 1, D has
    both signs (LL^t would fail) and the factors do not grow: two correct
    LDL^t's agree to the rounding of their operations."""
    n = int(np.prod(grid))
    L = problems.laplacian(*grid)
    sign = np.where(np.random.default_rng(seed).random(n) < 0.5, -1.0, 1.0)
    return (L - sp.diags(L.diagonal()) + sp.diags(7.0 * sign)).tocsc()


def _offdiag_mask(P):
    """Coefficients of the panels below the diagonal blocks."""
    mask = np.ones(P.S.coefnbr, dtype=bool)
    for k in range(P.S.cblknbr):
        f, l, fb, stride, off = P.cblk[k]
        nc = l - f + 1
        mask[off:off + nc * nc] = False
    return mask


class LdltProblem(Problem):
    def oracle_factors(self, dtype=np.float64):
        return orc.factorize(self.S, self.A, self.perm, 1, self.crit, dtype=dtype)

    def gpu_factors(self):
        return pa.get_coeftab(self.h, 0), pa.get_coeftab(self.h, 1)


@pytest.mark.parametrize("trsm", ["inverse", "substitution"])
@pytest.mark.parametrize("fact", [pa.factotype.LDLT, pa.factotype.LDLH], ids=["ldlt", "ldlh"])
def test_ldlt_factors_and_solve_match_the_oracle(fact, trsm, monkeypatch):
    """FP64 LDL^t of a symmetric indefinite matrix: the L array and the
    panel copy L D equal the oracle's within 1e-12, the solve matches a
    sparse direct solve.  "substitution" builds the schedule with the
    substitution TRSM and the separate L D copy kernel."""
    if trsm == "substitution":
        monkeypatch.setenv("PASTIX_AMD_TRSM", "0")
    grid = (12, 11, 10)
    print("case LDL^t", fact, trsm, flush=True)
    A = _indefinite_dd(grid, 1)
    P = LdltProblem(A, grid, fact)
    try:
        Lo, Do, npiv = P.oracle_factors()
        Lg, Dg = P.gpu_factors()
        assert npiv == P.iparm[pa.iparm.static_pivoting] == 0
        assert np.abs(Lg - Lo).max() / np.abs(Lo).max() <= 1e-12
        m = _offdiag_mask(P)
        assert np.abs(Dg[m] - Do[m]).max() / np.abs(Do[m]).max() <= 1e-12
        rng = np.random.default_rng(3)
        b = rng.uniform(-1, 1, (P.n, 2))
        x = pa.task_solve(P.h, P.spm, b.copy(order="F"))
        assert np.abs(x - spla.spsolve(A, b)).max() <= 1e-11 * np.abs(x).max()
        assert problems.backward_error(A, x[:, 0], b[:, 0]) <= 1e-14
        # the flop count of the LDL^t model
        th = float(P.dparm[pa.dparm.fact_thflops])
        assert th > 0
    finally:
        P.close()


def test_ldlt_solve_by_subtasks_with_the_diagonal_step():
    """applyorder + trsm(L, unit) + diag (D^{-1}) + trsm(L^t, unit) +
    applyorder gives pastix_task_solve's solution (reference
    sopalin/pastix_task_solve.c:596-640 for LDL^t)."""
    grid = (10, 9, 8)
    A = _indefinite_dd(grid, 2)
    P = LdltProblem(A, grid, pa.factotype.LDLT)
    lib = _lib.load()
    try:
        n = P.n
        rng = np.random.default_rng(4)
        b = np.asfortranarray(rng.uniform(-1, 1, (n, 2)))
        xref = pa.task_solve(P.h, P.spm, b.copy(order="F"))
        x = b.copy(order="F")
        Bp = C.c_void_p()
        assert lib.pastixRhsInit(C.byref(Bp)) == 0
        assert lib.pastix_subtask_applyorder(P.h, pa.dir.Forward, n, 2, C.c_void_p(x.ctypes.data), n, Bp) == 0
        assert lib.pastix_subtask_trsm(P.h, LEFT, LOWER, NOTRANS, UNIT, Bp) == 0
        assert lib.pastix_subtask_diag(P.h, Bp) == 0
        assert lib.pastix_subtask_trsm(P.h, LEFT, LOWER, TRANS, UNIT, Bp) == 0
        assert lib.pastix_subtask_applyorder(P.h, pa.dir.Backward, n, 2, C.c_void_p(x.ctypes.data), n, Bp) == 0
        assert np.abs(x - xref).max() <= 1e-13 * np.abs(xref).max()
        lib.pastixRhsFinalize(Bp)
    finally:
        P.close()


@pytest.mark.parametrize("refine", [pa.refine.GMRES, pa.refine.BiCGSTAB], ids=["gmres", "bicgstab"])
def test_ldlt_fp32_factors_refined_to_1e12(refine):
    """Mixed precision LDL^t: FP32 factors on the GPU equal the FP32
    oracle's within the FP32 rounding (diagonally dominant indefinite
    matrix), and FP64 refinement reaches backward error 1e-12 on it and on
    the shifted Laplacian, whose small pivots leave the FP32 solve far from
    the solution."""
    grid = (14, 13, 12)
    for name, A in (("dd", _indefinite_dd(grid, 3)), ("shifted", _indefinite(grid, 0.7))):
        print("case", name, flush=True)
        P = LdltProblem(A, grid, pa.factotype.LDLT, mixed=True)
        try:
            P.iparm[pa.iparm.refinement] = refine
            if name == "dd":
                Lg, Dg = P.gpu_factors()
                Lo, Do, npiv = P.oracle_factors(dtype=np.float32)
                assert np.abs(Lg - Lo).max() / np.abs(Lo).max() <= 1e-5
            b = A @ problems.uniform(P.n, seed=5)
            x = pa.task_solve(P.h, P.spm, b.copy())
            assert problems.backward_error(A, x, b) > 1e-10
            pa.task_refine(P.h, P.spm, b, x)
            assert problems.backward_error(A, x, b) <= 1e-12
        finally:
            P.close()
