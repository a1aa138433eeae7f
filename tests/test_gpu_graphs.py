"""The factorization and solve launch sequences replayed from hipGraphs give
the same results as the eager launches, on every replay, with unrelated
eager device work and synchronisations between the replays."""
import os

import numpy as np
import pytest

import pastix_amd as pa
from pastix_amd import problems

pytestmark = pytest.mark.gpu


def _make(A, grid, ft, mixed, graphs):
    old = os.environ.get("PASTIX_AMD_GRAPHS")
    os.environ["PASTIX_AMD_GRAPHS"] = "1" if graphs else "0"
    try:
        ip, dp = pa.initParam()
        ip[pa.iparm.factorization] = ft
        ip[pa.iparm.verbose] = pa.verbose.Not
        if mixed:
            ip[pa.iparm.mixed] = 1
        h = pa.init(ip, dp)
    finally:
        if old is None:
            del os.environ["PASTIX_AMD_GRAPHS"]
        else:
            os.environ["PASTIX_AMD_GRAPHS"] = old
    mt = pa.mtxtype.General if ft == pa.factotype.LU else pa.mtxtype.Symmetric
    spm = pa.spmatrix(A, mtxtype=mt)
    pa.task_analyze(h, spm, pa.order.grid(*grid))
    return h, spm


@pytest.mark.parametrize("ft,mixed", [(pa.factotype.LLT, False), (pa.factotype.LLT, True), (pa.factotype.LU, False)],
                         ids=["llt", "llt-mixed", "lu"])
def test_graph_replays_match_eager_launches(ft, mixed):
    import torch
    grid = (14, 13, 12)
    A = problems.convection_diffusion_27(*grid) if ft == pa.factotype.LU else problems.laplacian(*grid)
    n = A.shape[0]
    hg, sg = _make(A, grid, ft, mixed, True)
    he, se = _make(A, grid, ft, mixed, False)
    try:
        tol = 1e-13 if not mixed else 2e-6
        for rep in range(3):                      # the factorization graph replays
            pa.task_numfact(hg, sg)
            pa.task_numfact(he, se)
            Lg, Le = pa.get_coeftab(hg, 0), pa.get_coeftab(he, 0)
            assert np.abs(Lg - Le).max() <= tol * np.abs(Le).max(), rep
            junk = torch.randn(1 << 20, device="cuda")
            (junk * 2).sum().item()               # eager device work between replays
            torch.cuda.synchronize()
        rng = np.random.default_rng(8)
        for rep in range(4):                      # the solve graphs replay
            b = rng.uniform(-1, 1, n)
            xg, xe = b.copy(), b.copy()
            pa.task_solve(hg, sg, xg)
            pa.task_solve(he, se, xe)
            assert np.abs(xg - xe).max() <= (1e-12 if not mixed else 1e-5) * np.abs(xe).max(), rep
            torch.cuda.synchronize()
    finally:
        pa.finalize(hg)
        pa.finalize(he)
