"""Autotuner of the launch schedule (pastix_amd_tune): every candidate
schedule gives the same factors up to rounding (the choice of GEMM tile
classes and of the narrow-cblk kernels changes only the order of the
floating-point sums), the tuner keeps the fastest candidate, leaves a valid
factorization in place, and a chosen schedule can be applied again."""
import numpy as np
import pytest

import pastix_amd as pa
from pastix_amd import problems
from pastix_amd.solver import TUNE_CANDIDATES, TUNE_NCAND

pytestmark = pytest.mark.gpu


def _setup(A, grid, ft, mixed):
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = ft
    ip[pa.iparm.verbose] = pa.verbose.Not
    if mixed:
        ip[pa.iparm.mixed] = 1
    # ||r||_2/||b||_2 <= 1e-13 keeps the inf-norm backward error below 1e-12
    dp[pa.dparm.epsilon_refinement] = 1e-13
    h = pa.init(ip, dp)
    mt = pa.mtxtype.General if ft == pa.factotype.LU else pa.mtxtype.Symmetric
    spm = pa.spmatrix(A, mtxtype=mt)
    pa.task_analyze(h, spm, pa.order.grid(*grid))
    pa.task_numfact(h, spm)
    return h, spm


@pytest.mark.parametrize("ft,mixed", [(pa.factotype.LLT, True), (pa.factotype.LLT, False), (pa.factotype.LU, True)],
                         ids=["llt-fp32", "llt-fp64", "lu-fp32"])
def test_every_candidate_schedule_gives_the_same_factors(ft, mixed):
    grid = (26, 25, 24)
    A = problems.laplacian(*grid) if ft == pa.factotype.LLT else problems.convection_diffusion_27(*grid)
    h, spm = _setup(A, grid, ft, mixed)
    try:
        cblk, blok, perm, invp, dims = pa.get_structure(h)
        keep = np.ones(int(dims[3]), dtype=bool)
        if ft == pa.factotype.LLT:
            for f, l, fb, st, off in cblk:
                nc = l - f + 1
                keep[off:off + nc * nc].reshape((nc, nc))[np.tril_indices(nc, -1)] = False
        sides = (0, 1) if ft == pa.factotype.LU else (0,)
        ref = [pa.get_coeftab(h, s)[keep] for s in sides]
        tol = 2e-5 if mixed else 1e-13
        for params in TUNE_CANDIDATES:
            pa.set_tuning(h, params)
            pa.task_numfact(h, spm)
            for s, R in zip(sides, ref):
                L = pa.get_coeftab(h, s)[keep]
                assert np.abs(L - R).max() <= tol * np.abs(R).max(), (params, s)
        res = pa.tune(h)
        t = np.array(res["times"])
        assert 0 <= res["best"] < TUNE_NCAND
        assert t[res["best"]] > 0 and t[res["best"]] == t[t > 0].min()
        assert tuple(res["params"]) == TUNE_CANDIDATES[res["best"]]
        # the handle holds the chosen schedule's factorization: solve with it
        n = A.shape[0]
        xt = problems.uniform(n, seed=3)
        b = A @ xt
        x = pa.task_solve(h, spm, b.copy())
        if mixed:
            pa.task_refine(h, spm, b, x)
        assert problems.backward_error(A, x, b) <= 1e-12
    finally:
        pa.finalize(h)
