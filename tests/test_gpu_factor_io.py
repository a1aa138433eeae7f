"""Factor I/O on the GPU, through the C-ABI: binary checkpoints of the
factors in HBM (pastix_amd_factors_save / _load) and the reference's
human-readable dump (coeftab_ddump / coeftab_sdump, reference
sopalin/coeftab_z.c:55 and sopalin/coeftab_zinit.c:54).

A checkpoint loaded into a fresh instance (same analysis, matrix on the
device) holds the saved factors bit for bit and solves to the same solution
(within 1e-12 relative in FP64 and 1e-5 in FP32: the forward sweep adds its
updates with atomics, so two solves can differ in the last bits); a
mixed-precision checkpoint refines to 1e-12 on the loaded instance; a file
of another structure, factorization kind or precision, or a truncated one,
is refused.  The dump's lines are the nonzero factor entries at their (row,
column), equal to the factors read back from HBM."""
import ctypes as C
import glob
import os

import numpy as np
import pytest
import scipy.sparse as sp

import pastix_amd as pa
from pastix_amd import _lib, problems
from oracle import pyoracle as orc
from helpers import Problem

pytestmark = pytest.mark.gpu


def _fresh(A, grid, ft, mixed=False):
    """A new instance with the same parameters: analysis + spm2bcsc only."""
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = ft
    ip[pa.iparm.verbose] = pa.verbose.Not
    if mixed:
        ip[pa.iparm.mixed] = 1
    h = pa.init(ip, dp)
    mt = pa.mtxtype.General if ft == pa.factotype.LU else pa.mtxtype.Symmetric
    spm = pa.spmatrix(A, mtxtype=mt)
    pa.task_analyze(h, spm, pa.order.grid(*grid))
    assert _lib.load().pastix_subtask_spm2bcsc(h, spm.ref) == 0
    return h, spm, ip, dp


@pytest.mark.parametrize("case", ["llt", "lu-mixed", "ldlt"])
def test_checkpoint_round_trip_gives_the_same_factors_and_solution(case, tmp_path):
    ft = {"llt": pa.factotype.LLT, "lu-mixed": pa.factotype.LU, "ldlt": pa.factotype.LDLT}[case]
    mixed = case.endswith("mixed")
    grid = (10, 9, 8)
    gen = problems.convection_diffusion_27 if ft == pa.factotype.LU else problems.laplacian
    A = gen(*grid)
    print("checkpoint case", case, grid, flush=True)
    P = Problem(A, grid, ft, mixed=mixed)
    lib = _lib.load()
    path = str(tmp_path / "factors.bin").encode()
    try:
        assert lib.pastix_amd_factors_save(P.h, path) == 0
        L0, U0 = P.gpu_factors() if ft == pa.factotype.LU else (pa.get_coeftab(P.h, 0), None)
        rng = np.random.default_rng(3)
        b = rng.uniform(-1, 1, P.n)
        x0 = pa.task_solve(P.h, P.spm, b.copy())
        h, spm, ip, dp = _fresh(A, grid, ft, mixed)
        try:
            assert lib.pastix_amd_factors_load(h, path) == 0
            L1 = pa.get_coeftab(h, 0)
            assert np.array_equal(L1, L0)
            if U0 is not None:
                assert np.array_equal(pa.get_coeftab(h, 1), U0)
            x1 = pa.task_solve(h, spm, b.copy())
            tol = 1e-5 if mixed else 1e-12
            assert np.abs(x1 - x0).max() <= tol * np.abs(x0).max()
            if mixed:
                x = x1.copy()
                pa.task_refine(h, spm, b.copy(), x)
                assert problems.backward_error(A, x, b) <= 1e-12
        finally:
            pa.finalize(h)
    finally:
        P.close()


def test_checkpoint_of_another_structure_kind_or_precision_is_refused(tmp_path):
    grid = (8, 8, 8)
    A = problems.laplacian(*grid)
    P = Problem(A, grid, pa.factotype.LLT)
    lib = _lib.load()
    path = str(tmp_path / "f.bin").encode()
    try:
        assert lib.pastix_amd_factors_save(P.h, path) == 0
        # another grid
        A2 = problems.laplacian(8, 8, 7)
        h, spm, _, _ = _fresh(A2, (8, 8, 7), pa.factotype.LLT)
        assert lib.pastix_amd_factors_load(h, path) == pa.error.BADPARAMETER
        pa.finalize(h)
        # another factorization kind on the same structure
        h, spm, _, _ = _fresh(A, grid, pa.factotype.LDLT)
        assert lib.pastix_amd_factors_load(h, path) == pa.error.BADPARAMETER
        pa.finalize(h)
        # another precision
        h, spm, _, _ = _fresh(A, grid, pa.factotype.LLT, mixed=True)
        assert lib.pastix_amd_factors_load(h, path) == pa.error.BADPARAMETER
        pa.finalize(h)
        # truncated
        raw = open(path, "rb").read()
        tp = str(tmp_path / "t.bin")
        open(tp, "wb").write(raw[:len(raw) // 2])
        h, spm, _, _ = _fresh(A, grid, pa.factotype.LLT)
        assert lib.pastix_amd_factors_load(h, tp.encode()) == pa.error.FILE
        # the refused load left no factorization behind
        x = np.ones(P.n)
        with pytest.raises(pa.PastixError):
            pa.task_solve(h, spm, x)
        pa.finalize(h)
    finally:
        P.close()


def _parse_dump(prefix):
    ent = {}
    for fn in glob.glob(prefix + "_*.txt"):
        for line in open(fn):
            i, j, v, side = line.split()
            ent[(side, int(i), int(j))] = float(v)
    return ent


@pytest.mark.parametrize("lu", [False, True], ids=["llt", "lu"])
def test_coeftab_dump_lists_the_factor_entries(lu, tmp_path):
    grid = (6, 5, 4)
    A = problems.convection_diffusion_27(*grid) if lu else problems.laplacian(*grid)
    P = Problem(A, grid, pa.factotype.LU if lu else pa.factotype.LLT)
    lib = _lib.load()
    try:
        solv = lib.pastix_amd_get_solver(P.h)
        assert solv
        prefix = str(tmp_path / "potrf")
        lib.coeftab_ddump(P.h, C.c_void_p(solv), prefix.encode())
        ent = _parse_dump(prefix)
        assert len(glob.glob(prefix + "_*.txt")) == P.S.cblknbr
        n = P.n
        L = np.tril(orc.expand_L(P.S, pa.get_coeftab(P.h, 0), n))
        nz = np.argwhere(L != 0)
        Lkeys = [k for k in ent if k[0] == "[L]"]
        assert len(Lkeys) == len(nz)
        for r, c in nz:
            # "%13e": 7 significant digits
            assert abs(ent[("[L]", r, c)] - L[r, c]) <= 1e-6 * abs(L[r, c])
        if lu:
            U = np.tril(orc.expand_L(P.S, pa.get_coeftab(P.h, 1), n))
            Ukeys = [k for k in ent if k[0] == "[U]"]
            assert len(Ukeys) == np.count_nonzero(U)
            for r, c in np.argwhere(U != 0):
                assert abs(ent[("[U]", c, r)] - U[r, c]) <= 1e-6 * abs(U[r, c])
        # the single precision dump refuses FP64 factors and writes nothing
        p2 = str(tmp_path / "s")
        lib.coeftab_sdump(P.h, C.c_void_p(solv), p2.encode())
        assert glob.glob(p2 + "_*.txt") == []
    finally:
        P.close()
