"""The allocation-free solve entry point (pastix_amd_solve_dev_noalloc):
through the C-ABI, its solution equals pastix_amd_solve_dev's on the same
device right-hand sides, on two calls in a row the second of which is
smaller, for FP64 and FP32 factors; a call beyond the scratch limit is
refused before any launch and leaves the right-hand sides untouched.

Both entry points run the same kernels, but the forward sweep's updates are
added to the right-hand sides with atomics, so the order of those additions
(and the last bits of the result) can differ between two runs of either
entry point: the solutions agree within 1e-12 relative (FP64, the factor
tolerance of the parity tests) and 1e-5 relative (FP32 sweeps)."""
import ctypes as C

import numpy as np
import pytest
import torch

import pastix_amd as pa
from pastix_amd import _lib, problems
from helpers import Problem

pytestmark = pytest.mark.gpu

NOTRANS, TRANS = 111, 112


def _solve_both(P, b, trans):
    """(solution of pastix_amd_solve_dev, solution of the no-alloc entry)."""
    lib = _lib.load()
    n, nrhs = b.shape
    ref = torch.tensor(b.T.copy(), dtype=torch.float64, device="cuda:0")  # column-major rows
    out = ref.clone()
    torch.cuda.synchronize()
    assert lib.pastix_amd_solve_dev(P.h, trans, nrhs, C.c_void_p(ref.data_ptr()), n) == 0
    assert lib.pastix_amd_solve_dev_noalloc(P.h, trans, nrhs, C.c_void_p(out.data_ptr()), n) == 0
    torch.cuda.synchronize()
    return ref.cpu().numpy().T, out.cpu().numpy().T


@pytest.mark.parametrize("mixed", [False, True])
def test_noalloc_solve_equals_the_allocating_solve_twice_the_second_smaller(mixed):
    rng = np.random.default_rng(5)
    cases = [(pa.factotype.LU, (11, 10, 9), problems.convection_diffusion_27, 3, TRANS),
             (pa.factotype.LLT, (7, 6, 5), problems.laplacian, 1, NOTRANS)]
    for ft, grid, gen, nrhs, trans in cases:
        print("noalloc case", ft, grid, nrhs, "mixed", mixed, flush=True)
        A = gen(*grid)
        P = Problem(A, grid, ft, mixed=mixed)
        try:
            b = rng.uniform(-1, 1, (P.n, nrhs))
            xr, xn = _solve_both(P, b, trans)
            tol = 1e-5 if mixed else 1e-12
            assert np.abs(xn - xr).max() <= tol * np.abs(xr).max()
            Aop = A.T if trans == TRANS else A
            r = np.abs(Aop @ xn - b).max() / np.abs(b).max()
            assert r < (1e-4 if mixed else 1e-12)
        finally:
            P.close()


def test_noalloc_solve_refuses_calls_beyond_the_scratch():
    grid = (8, 8, 8)
    P = Problem(problems.laplacian(*grid), grid, pa.factotype.LLT)
    lib = _lib.load()
    try:
        nrhs = (128 << 20) // (16 * P.n) + 1   # n * nrhs * 16 bytes > 128 MiB
        b = torch.ones((nrhs, P.n), dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        rc = lib.pastix_amd_solve_dev_noalloc(P.h, NOTRANS, nrhs, C.c_void_p(b.data_ptr()), P.n)
        assert rc == pa.error.BADPARAMETER
        torch.cuda.synchronize()
        assert bool((b == 1).all())
    finally:
        P.close()
