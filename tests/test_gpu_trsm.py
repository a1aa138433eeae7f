"""The TRSM by inverted 16 x 16 diagonal tiles (k_tile_inv + k_trsm_inv, the
default) gives the factors of the substitution TRSM (k_trsm, selected with
PASTIX_AMD_TRSM=0), for LL^t and LU in FP64 and FP32, on grids whose cblk
widths are not multiples of 16 and whose row slabs end short of 32 rows; the
FP64 factors also match the CPU oracle."""
import os

import numpy as np
import pytest

import pastix_amd as pa
from pastix_amd import problems
from helpers import Problem

pytestmark = pytest.mark.gpu


def _factors(A, grid, ft, mixed, inv, bs):
    old = os.environ.get("PASTIX_AMD_TRSM")
    os.environ["PASTIX_AMD_TRSM"] = "1" if inv else "0"
    try:
        P = Problem(A, grid, ft, mixed=mixed, blocksize=bs)
        L, U = P.gpu_factors()
    finally:
        if old is None:
            del os.environ["PASTIX_AMD_TRSM"]
        else:
            os.environ["PASTIX_AMD_TRSM"] = old
    return P, L, U


CASES = [
    ("llt-fp64", pa.factotype.LLT, False, (17, 15, 13), None),
    ("llt-fp64-bs", pa.factotype.LLT, False, (12, 11, 10), (20, 40)),
    ("llt-fp32", pa.factotype.LLT, True, (17, 15, 13), None),
    ("lu-fp64", pa.factotype.LU, False, (13, 12, 11), None),
    ("lu-fp64-bs", pa.factotype.LU, False, (10, 9, 8), (20, 40)),
    ("lu-fp32", pa.factotype.LU, True, (13, 12, 11), None),
]


@pytest.mark.parametrize("name,ft,mixed,grid,bs", CASES, ids=[c[0] for c in CASES])
def test_inverse_tile_trsm_matches_substitution(name, ft, mixed, grid, bs):
    print("case", name, grid, bs, flush=True)
    lu = ft == pa.factotype.LU
    A = problems.convection_diffusion_27(*grid) if lu else problems.laplacian(*grid)
    P1, L1, U1 = _factors(A, grid, ft, mixed, True, bs)
    P0, L0, U0 = _factors(A, grid, ft, mixed, False, bs)
    try:
        ncol = P1.cblk[:, 1] - P1.cblk[:, 0] + 1
        assert (ncol % 16 != 0).any()              # partial last column tiles
        mask = None if lu else P1.lower_mask()
        tol = 2e-5 if mixed else 1e-12
        for G, H in ((L1, L0),) + (((U1, U0),) if lu else ()):
            g, h = (G, H) if mask is None else (G[mask], H[mask])
            assert np.abs(g - h).max() <= tol * np.abs(h).max()
        if not mixed:
            Lo, Uo, _ = P1.oracle_factors()
            g, o = (L1, Lo) if mask is None else (L1[mask], Lo[mask])
            assert np.abs(g - o).max() <= 1e-12 * np.abs(o).max()
            if lu:
                assert np.abs(U1 - Uo).max() <= 1e-12 * np.abs(Uo).max()
        # the solve with the new factors
        n = P1.n
        b = problems.uniform(n, seed=4)
        x = pa.task_solve(P1.h, P1.spm, b.copy())
        assert problems.backward_error(A, x, b) <= (1e-5 if mixed else 1e-14)
    finally:
        P1.close()
        P0.close()
