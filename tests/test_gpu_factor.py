"""GPU parity of the level-batched factorization and solve against the CPU
oracle on the same supernodal structure (FP64 within 1e-12 relative, FP32
within a float tolerance), and end-to-end accuracy of solve + refinement.
"""
import numpy as np
import pytest

import pastix_amd as pa
from pastix_amd import problems
from oracle import pyoracle as orc
from helpers import Problem

pytestmark = pytest.mark.gpu

CASES_LLT = [
    ("lap3d_6", (6, 6, 6), None),
    ("lap3d_9x7x5_bs", (9, 7, 5), (4, 8)),
    ("lap3d_20", (20, 20, 20), None),
    ("lap2d_40x33", (40, 33, 1), (8, 16)),
]


def _relerr(a, b, mask=None):
    if mask is not None:
        a, b = a[mask], b[mask]
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("name,grid,bs", CASES_LLT, ids=[c[0] for c in CASES_LLT])
def test_potrf_fp64_matches_oracle(name, grid, bs):
    print("case", name, grid, bs, flush=True)
    A = problems.laplacian(*grid)
    P = Problem(A, grid, pa.factotype.LLT, blocksize=bs)
    try:
        Lg, _ = P.gpu_factors()
        Lo, _, npiv = P.oracle_factors()
        mask = P.lower_mask()
        assert _relerr(Lg, Lo, mask) < 1e-12
        assert P.iparm[pa.iparm.static_pivoting] == npiv == 0
    finally:
        P.close()


CASES_LU = [
    ("cd27_5", (5, 5, 5), None),
    ("cd27_9x8x7_bs", (9, 8, 7), (6, 12)),
    ("cd27_14", (14, 14, 14), None),
]


@pytest.mark.parametrize("name,grid,bs", CASES_LU, ids=[c[0] for c in CASES_LU])
def test_getrf_fp64_matches_oracle(name, grid, bs):
    print("case", name, grid, bs, flush=True)
    A = problems.convection_diffusion_27(*grid)
    P = Problem(A, grid, pa.factotype.LU, blocksize=bs)
    try:
        Lg, Ug = P.gpu_factors()
        Lo, Uo, npiv = P.oracle_factors()
        assert _relerr(Lg, Lo) < 1e-12
        assert _relerr(Ug, Uo) < 1e-12
    finally:
        P.close()


@pytest.mark.parametrize("ft,grid", [(pa.factotype.LLT, (10, 9, 8)), (pa.factotype.LU, (10, 9, 8)),
                                     (pa.factotype.LLT, (26, 25, 24)), (pa.factotype.LU, (19, 18, 17))],
                         ids=["llt", "lu", "llt-128row-tiles", "lu-128row-tiles"])
def test_fp32_factor_matches_oracle(ft, grid):
    """The larger grids have updates of >= 384 rows, which take the 128-row
    GEMM tiles (128 x 128, 128 x 64 and 128 x 32)."""
    A = problems.laplacian(*grid) if ft == pa.factotype.LLT else problems.convection_diffusion_27(*grid)
    P = Problem(A, grid, ft, mixed=True)
    try:
        Lg, Ug = P.gpu_factors()
        Lo, Uo, _ = P.oracle_factors(dtype=np.float32)
        mask = None if P.lu else P.lower_mask()
        # same algorithm in FP32, different summation order: a few ulps times
        # the growth of the elimination
        assert _relerr(Lg, Lo.astype(np.float64), mask) < 2e-5
        if P.lu:
            assert _relerr(Ug, Uo.astype(np.float64)) < 2e-5
    finally:
        P.close()


@pytest.mark.parametrize("ft", [pa.factotype.LLT, pa.factotype.LU], ids=["llt", "lu"])
def test_solve_matches_oracle_and_is_accurate(ft):
    grid = (12, 11, 10)
    A = problems.laplacian(*grid) if ft == pa.factotype.LLT else problems.convection_diffusion_27(*grid)
    P = Problem(A, grid, ft)
    try:
        n = P.n
        xt = problems.uniform(n, seed=0)
        b = A @ xt
        x = pa.task_solve(P.h, P.spm, b.copy())
        Lo, Uo, _ = P.oracle_factors()
        xo_p = orc.solve(P.S, 2 if P.lu else 0, Lo, Uo, b[P.invp])
        xo = np.empty(n)
        xo[P.invp] = xo_p
        assert np.abs(x - xo).max() / np.abs(xo).max() < 1e-12
        assert problems.backward_error(A, x, b) < 1e-14
    finally:
        P.close()


def test_multi_rhs_solve():
    grid = (10, 10, 10)
    A = problems.laplacian(*grid)
    P = Problem(A, grid, pa.factotype.LLT)
    try:
        X = problems.uniform(P.n, seed=1, nrhs=16)
        B = np.asfortranarray(A @ X)
        Y = pa.task_solve(P.h, P.spm, B.copy(order="F"))
        assert np.abs(Y - X).max() < 1e-11
    finally:
        P.close()


@pytest.mark.parametrize("refine", [pa.refine.GMRES, pa.refine.CG, pa.refine.SR], ids=["gmres", "cg", "sr"])
def test_mixed_precision_refinement_reaches_1e12(refine):
    grid = (16, 16, 16)
    A = problems.laplacian(*grid)
    P = Problem(A, grid, pa.factotype.LLT, mixed=True)
    try:
        P.iparm[pa.iparm.refinement] = refine
        xt = problems.uniform(P.n, seed=0)
        b = A @ xt
        x = pa.task_solve(P.h, P.spm, b.copy())
        berr0 = problems.backward_error(A, x, b)
        assert berr0 > 1e-10  # FP32 factors alone do not reach 1e-12
        pa.task_refine(P.h, P.spm, b, x)
        # stopping rule ||r||_2/||b||_2 <= 1e-12 (reference refinement/z_refine_gmres.c:169)
        assert P.dparm[pa.dparm.relative_error] <= 1e-12
        # normwise backward error ||b-Ax||_inf / (||A||_inf ||x||_inf + ||b||_inf)
        assert problems.backward_error(A, x, b) <= 1e-12
        assert P.iparm[pa.iparm.nbiter] >= 1
    finally:
        P.close()


def test_mixed_lu_gmres_convection_diffusion():
    grid = (14, 13, 12)
    A = problems.convection_diffusion_27(*grid)
    P = Problem(A, grid, pa.factotype.LU, mixed=True)
    try:
        xt = problems.uniform(P.n, seed=0)
        b = A @ xt
        x = pa.task_solve(P.h, P.spm, b.copy())
        pa.task_refine(P.h, P.spm, b, x)
        assert P.dparm[pa.dparm.relative_error] <= 1e-12
        assert problems.backward_error(A, x, b) <= 1e-12
    finally:
        P.close()


def test_static_pivoting_counts_replaced_pivots():
    """The first pivot of the factorization is exactly zero (its row and
    column are empty): it is replaced by the threshold eps*||A||_F and counted
    (reference kernels/core_zgetrfsp.c:81-91), on the GPU and in the oracle."""
    grid = (5, 4, 3)
    A = problems.convection_diffusion_27(*grid)
    P0 = Problem(A, grid, pa.factotype.LU, analyze_only=True)
    v = int(P0.invp[0])
    P0.close()
    A = A.tolil()
    A[v, :] = 0.0
    A[:, v] = 0.0
    A = A.tocsc()
    A.eliminate_zeros()
    P = Problem(A, grid, pa.factotype.LU)
    try:
        Lg, Ug = P.gpu_factors()
        Lo, Uo, npiv = P.oracle_factors()
        assert npiv == 1
        assert P.iparm[pa.iparm.static_pivoting] == npiv
        assert _relerr(Lg, Lo) < 1e-12
        assert _relerr(Ug, Uo) < 1e-12
    finally:
        P.close()


@pytest.mark.parametrize("ft", [pa.factotype.LLT, pa.factotype.LU], ids=["llt", "lu"])
def test_fp64_factors_match_oracle_on_graded_anisotropic_coefficients(ft):
    """Cell conductivities exp(N(0,1)) with a 10x z-anisotropy: the diagonal
    tiles the TRSM inverts span several orders of magnitude, and the factors
    (TRSM by inverted 16 x 16 diagonal tiles) still equal the oracle's
    (substitution, as the reference kernel) within 1e-12."""
    grid = (17, 15, 13)
    A = problems.anisotropic_diffusion(*grid, problems.anisotropic_conductivity(*grid, seed=7))
    print("case anisotropic", grid, ft, flush=True)
    P = Problem(A, grid, ft)
    try:
        Lg, Ug = P.gpu_factors()
        Lo, Uo, _ = P.oracle_factors()
        mask = None if P.lu else P.lower_mask()
        assert _relerr(Lg, Lo, mask) < 1e-12
        if P.lu:
            assert _relerr(Ug, Uo) < 1e-12
        b = problems.uniform(P.n, seed=3)
        x = pa.task_solve(P.h, P.spm, b.copy())
        assert problems.backward_error(A, x, b) < 1e-14
    finally:
        P.close()



def test_orsirr_fixture_with_a_personal_ordering():
    """The reference's LU test matrix (test/matrix/orsirr.rua, n = 1030) with
    a reverse Cuthill-McKee permutation given as a bare personal ordering:
    the analysis finds and amalgamates the supernodes (reference
    order/pastix_subtask_order.c:362-381); the FP64 LU factors equal the
    oracle's on that structure, and an FP32 factorization refined by FP64
    GMRES reaches backward error 1e-12."""
    import os
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    from test_cpu_oracle import read_rua
    A = read_rua(os.path.join(os.path.dirname(__file__), "golden", "orsirr.rua"))
    pat = (abs(A) + abs(A.T)).tocsr()
    rcm = reverse_cuthill_mckee(pat, symmetric_mode=True).astype(np.int64)
    perm = np.empty_like(rcm)
    perm[rcm] = np.arange(len(rcm))
    print("case orsirr LU FP64", flush=True)
    P = Problem(A, None, pa.factotype.LU, order=pa.order.from_permutation(perm))
    try:
        assert P.S.cblknbr < P.n // 4   # real supernodes, not one column each
        Lo, Uo, npiv = P.oracle_factors()
        Lg, Ug = P.gpu_factors()
        assert npiv == P.iparm[pa.iparm.static_pivoting]
        assert np.abs(Lg - Lo).max() / np.abs(Lo).max() <= 1e-12
        assert np.abs(Ug - Uo).max() / np.abs(Uo).max() <= 1e-12
        b = A @ problems.uniform(P.n, seed=0)
        x = pa.task_solve(P.h, P.spm, b.copy())
        assert problems.backward_error(A, x, b) <= 1e-14
    finally:
        P.close()
    print("case orsirr LU FP32 + GMRES", flush=True)
    P = Problem(A, None, pa.factotype.LU, mixed=True, order=pa.order.from_permutation(perm))
    try:
        x = pa.task_solve(P.h, P.spm, b.copy())
        pa.task_refine(P.h, P.spm, b, x)
        assert problems.backward_error(A, x, b) <= 1e-12
    finally:
        P.close()


@pytest.mark.parametrize("ft,mixed", [(pa.factotype.LLT, False), (pa.factotype.LU, False), (pa.factotype.LLT, True),
                                      (pa.factotype.LDLT, False)], ids=["llt", "lu", "llt-fp32", "ldlt"])
def test_lookahead_gives_the_factors_of_the_level_by_level_schedule(ft, mixed, monkeypatch):
    """With look-ahead (the default) the next level's diagonal blocks and
    TRSMs run on a second stream while the current level's other GEMMs run;
    the factors equal those of the level-by-level order
    (PASTIX_AMD_LOOKAHEAD=0) within the rounding of reordered atomic
    updates (1e-13 FP64, 1e-6 FP32), through the captured graph and eager."""
    grid = (22, 20, 18)
    gen = problems.convection_diffusion_27 if ft == pa.factotype.LU else problems.laplacian
    A = gen(*grid)
    if ft == pa.factotype.LDLT:
        # symmetric indefinite, strictly diagonally dominant (+-7 diagonal):
        # no pivot growth to amplify the reordered rounding
        import scipy.sparse as sp
        sign = np.where(np.random.default_rng(4).random(A.shape[0]) < 0.5, -1.0, 1.0)
        A = (A - sp.diags(A.diagonal()) + sp.diags(7.0 * sign)).tocsc()
    out = {}
    for mode, env in (("seq", {"PASTIX_AMD_LOOKAHEAD": "0"}), ("la", {}),
                      ("la-eager", {"PASTIX_AMD_GRAPHS": "0"})):
        for k in ("PASTIX_AMD_LOOKAHEAD", "PASTIX_AMD_GRAPHS"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        print("case lookahead", ft, mixed, mode, flush=True)
        P = Problem(A, grid, ft, mixed=mixed)
        try:
            # a second factorization replays the captured launch sequence
            pa.task_numfact(P.h, P.spm)
            out[mode] = P.gpu_factors()
            mask = None if P.lu or ft == pa.factotype.LDLT else P.lower_mask()
        finally:
            P.close()
    tol = 1e-6 if mixed else 1e-13
    for mode in ("la", "la-eager"):
        assert _relerr(out[mode][0], out["seq"][0], mask) < tol
        if out["seq"][1] is not None:
            assert _relerr(out[mode][1], out["seq"][1]) < tol


@pytest.mark.parametrize("ft,trans", [(pa.factotype.LLT, False), (pa.factotype.LU, False), (pa.factotype.LU, True)],
                         ids=["llt", "lu", "lu-trans"])
def test_solve_by_inverted_diagonal_blocks_matches_substitution(ft, trans, monkeypatch):
    """The diagonal-block solves by products with the inverted blocks that
    the TRSM launches write (the default) give the solution of the
    substitution kernels (PASTIX_AMD_DINV=0) within 1e-12, for both sweeps
    and both factor arrays; also after the factors were saved and loaded
    into a fresh instance (the load rebuilds the inverses)."""
    import os
    import tempfile
    grid = (21, 19, 17)
    gen = problems.convection_diffusion_27 if ft == pa.factotype.LU else problems.laplacian
    A = gen(*grid)
    rng = np.random.default_rng(8)
    b = np.asfortranarray(rng.uniform(-1, 1, (A.shape[0], 3)))
    xs = {}
    for mode in ("subst", "dinv"):
        monkeypatch.delenv("PASTIX_AMD_DINV", raising=False)
        if mode == "subst":
            monkeypatch.setenv("PASTIX_AMD_DINV", "0")
        print("case dinv", ft, trans, mode, flush=True)
        P = Problem(A, grid, ft)
        try:
            P.iparm[pa.iparm.transpose_solve] = pa.trans.Trans if trans else pa.trans.NoTrans
            xs[mode] = pa.task_solve(P.h, P.spm, b.copy(order="F"))
            if mode == "dinv":
                path = os.path.join(tempfile.mkdtemp(), "f.bin").encode()
                from pastix_amd import _lib
                assert _lib.load().pastix_amd_factors_save(P.h, path) == 0
                ip, dp = pa.initParam()
                ip[pa.iparm.factorization] = ft
                ip[pa.iparm.verbose] = pa.verbose.Not
                ip[pa.iparm.transpose_solve] = P.iparm[pa.iparm.transpose_solve]
                h2 = pa.init(ip, dp)
                mt = pa.mtxtype.General if ft == pa.factotype.LU else pa.mtxtype.Symmetric
                spm2 = pa.spmatrix(A, mtxtype=mt)
                pa.task_analyze(h2, spm2, pa.order.grid(*grid))
                assert _lib.load().pastix_subtask_spm2bcsc(h2, spm2.ref) == 0
                assert _lib.load().pastix_amd_factors_load(h2, path) == 0
                xs["loaded"] = pa.task_solve(h2, spm2, b.copy(order="F"))
                pa.finalize(h2)
        finally:
            P.close()
    Aop = A.T if trans else A
    for mode in ("dinv", "loaded"):
        assert np.abs(xs[mode] - xs["subst"]).max() <= 1e-12 * np.abs(xs["subst"]).max()
        for r in range(3):
            assert problems.backward_error(Aop, xs[mode][:, r], b[:, r]) <= 1e-14


@pytest.mark.parametrize("ft", [pa.factotype.LLT, pa.factotype.LU], ids=["llt", "lu"])
def test_performed_flops_count_the_padded_gemm_tiles(ft):
    """DPARM_FACT_RLFLOPS counts the flops the kernels perform (the reference
    accumulates its kernels' own counts, kernels/kernels_trace.c:349): the
    GEMM tiles padded to their class, so at least the model's THFLOPS and
    less than twice it on a 3D stencil."""
    grid = (20, 20, 20)
    gen = problems.convection_diffusion_27 if ft == pa.factotype.LU else problems.laplacian
    P = Problem(gen(*grid), grid, ft)
    try:
        th = float(P.dparm[pa.dparm.fact_thflops])
        rl = float(P.dparm[pa.dparm.fact_rlflops])
        assert th > 0 and th <= rl < 2 * th
    finally:
        P.close()
