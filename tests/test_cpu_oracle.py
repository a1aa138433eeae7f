"""Pinning of the CPU oracle (oracle/oracle.c) against what the mathematics
fixes: LAPACK/numpy Cholesky, a brute-force no-pivot Doolittle LU, scipy's
triangular solves and sparse direct solve, the closed form of the static
pivot, and the reference's own input fixture (test/matrix/orsirr.rua).
These run on the CPU only."""
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import pastix_amd as pa
from pastix_amd import problems
from oracle import pyoracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))


def doolittle(A):
    """No-pivot LU by the textbook triple loop (unit lower L, upper U)."""
    n = A.shape[0]
    U = np.array(A, dtype=np.float64)
    L = np.eye(n)
    for k in range(n):
        L[k + 1:, k] = U[k + 1:, k] / U[k, k]
        U[k + 1:, k:] -= np.outer(L[k + 1:, k], U[k, k:])
        U[k + 1:, k] = 0.
    return L, U


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("nc", [1, 5, 63, 64, 65, 130])
def test_potrf_diag_is_lapack_cholesky(nc):
    rng = np.random.default_rng(nc)
    M = rng.uniform(-1, 1, (nc, nc))
    M = M @ M.T + nc * np.eye(nc)
    D = np.asfortranarray(np.tril(M)).ravel(order="F").copy()
    npiv = np.zeros(1, dtype=np.int64)
    nneg = np.zeros(1, dtype=np.int64)
    orc.load().oracle_dpotrf_diag(nc, orc._p(D), 1e-300, orc._p(npiv), orc._p(nneg))
    L = np.tril(D.reshape((nc, nc), order="F"))
    assert _rel(L, np.linalg.cholesky(M)) < 1e-13
    assert npiv[0] == 0 and nneg[0] == 0


def test_potrf_static_pivot_closed_form():
    """A pivot below the criterion is replaced by the criterion and counted
    (reference kernels/core_zpotrfsp.c:80-84): for diag(1e-20, 4) with
    crit 1e-10 the factor is diag(sqrt(1e-10), 2) with one pivot."""
    D = np.array([1e-20, 0., 0., 4.])
    npiv = np.zeros(1, dtype=np.int64)
    nneg = np.zeros(1, dtype=np.int64)
    orc.load().oracle_dpotrf_diag(2, orc._p(D), 1e-10, orc._p(npiv), orc._p(nneg))
    assert np.isclose(D[0], np.sqrt(1e-10), rtol=1e-15) and D[3] == 2.
    assert npiv[0] == 1


@pytest.mark.parametrize("nc", [1, 7, 64, 65, 100])
def test_getrf_diag_is_doolittle(nc):
    rng = np.random.default_rng(100 + nc)
    M = rng.uniform(-1, 1, (nc, nc)) + nc * np.eye(nc)
    Lt = np.asfortranarray(np.tril(M)).ravel(order="F").copy()
    Ut = np.asfortranarray(np.tril(M.T, -1)).ravel(order="F").copy()
    npiv = np.zeros(1, dtype=np.int64)
    orc.load().oracle_dgetrf_diag(nc, orc._p(Lt), orc._p(Ut), 1e-300, orc._p(npiv))
    F = Lt.reshape((nc, nc), order="F")
    Lr, Ur = doolittle(M)
    assert _rel(np.tril(F, -1), np.tril(Lr, -1)) < 1e-12
    assert _rel(np.triu(F), Ur) < 1e-12
    # the U^T array holds the transposed factored square
    assert _rel(Ut.reshape((nc, nc), order="F"), F.T) == 0.


def _analyse(A, grid, lu, blocksize=(6, 12)):
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = pa.factotype.LU if lu else pa.factotype.LLT
    ip[pa.iparm.verbose] = pa.verbose.Not
    ip[pa.iparm.min_blocksize], ip[pa.iparm.max_blocksize] = blocksize
    ip[pa.iparm.tasks2d_width] = blocksize[0]
    h = pa.init(ip, dp)
    try:
        spm = pa.spmatrix(A, mtxtype=pa.mtxtype.General if lu else pa.mtxtype.Symmetric)
        pa.task_analyze(h, spm, pa.order.grid(*grid))
        cblk, blok, perm, invp, dims = pa.get_structure(h)
        th = float(dp[pa.dparm.fact_thflops])
    finally:
        pa.finalize(h)
    return orc.Structure(cblk, blok), perm, invp, th


def _expand(S, F, n):
    """Dense copy of a coeftab (all tiles, both triangles of the diagonal)."""
    M = np.zeros((n, n))
    for k in range(S.cblknbr):
        f, l, fb, stride, off = S.cblk[k]
        nc = l - f + 1
        for b in range(fb, S.cblkx[k + 1, 2]):
            fr, lr, _, _, coef = S.blok[b]
            m = lr - fr + 1
            M[fr:lr + 1, f:l + 1] = F[off + coef: off + coef + m * nc].reshape((nc, m)).T
    return M


def test_llt_factorization_is_the_unique_cholesky():
    grid = (6, 5, 4)
    A = problems.laplacian(*grid)
    S, perm, invp, th = _analyse(A, grid, lu=False)
    L, _, npiv = orc.factorize(S, A, perm, 0, 1e-300)
    n = A.shape[0]
    Ap = A.toarray()[np.ix_(invp, invp)]  # Ap[new i, new j] = A[old, old]
    Ld = np.tril(_expand(S, L, n))
    assert _rel(Ld, np.linalg.cholesky(Ap)) < 1e-13
    assert npiv == 0


@pytest.mark.parametrize("lu,use_blas", [(False, False), (False, True), (True, False), (True, True)],
                         ids=["llt", "llt-blas", "lu", "lu-blas"])
def test_parallel_oracle_matches_the_sequential_one(lu, use_blas):
    """The level-parallel drivers of the CPU baseline (4 threads, optional
    BLAS kernels) compute the factors of the sequential oracle, which the
    tests above pin to LAPACK / Doolittle."""
    grid = (12, 11, 10)
    A = problems.convection_diffusion_27(*grid) if lu else problems.laplacian(*grid)
    S, perm, invp, th = _analyse(A, grid, lu=lu)
    crit = np.finfo(np.float64).eps / 2 * sp.linalg.norm(A)
    L, U, npiv = orc.factorize(S, A, perm, 2 if lu else 0, crit)
    Lp, Up, npivp, sec = orc.factorize_parallel(S, A, perm, crit, 4, use_blas=use_blas, factotype=2 if lu else 0)
    n = A.shape[0]
    Ld, Lpd = _expand(S, L, n), _expand(S, Lp, n)
    if lu:
        assert _rel(Lpd, Ld) < 1e-12
        assert _rel(_expand(S, Up, n), _expand(S, U, n)) < 1e-12
    else:
        assert _rel(np.tril(Lpd), np.tril(Ld)) < 1e-12
    assert npivp == npiv and sec > 0


def test_lu_factorization_is_the_unique_doolittle():
    grid = (5, 4, 3)
    A = problems.convection_diffusion_27(*grid)
    S, perm, invp, th = _analyse(A, grid, lu=True)
    L, U, npiv = orc.factorize(S, A, perm, 2, 1e-300)
    n = A.shape[0]
    Ap = A.toarray()[np.ix_(invp, invp)]
    Lr, Ur = doolittle(Ap)
    Ld = _expand(S, L, n)
    Ud = _expand(S, U, n).T
    assert _rel(np.tril(Ld, -1), np.tril(Lr, -1)) < 1e-12
    # U lives transposed in the second array (its diagonal squares hold L + U)
    assert _rel(np.triu(Ud), Ur) < 1e-12


@pytest.mark.parametrize("lu", [False, True], ids=["llt", "lu"])
def test_oracle_solve_matches_sparse_direct_solve(lu):
    grid = (6, 5, 4)
    A = problems.convection_diffusion_27(*grid) if lu else problems.laplacian(*grid)
    S, perm, invp, th = _analyse(A, grid, lu=lu)
    L, U, _ = orc.factorize(S, A, perm, 2 if lu else 0, 1e-300)
    n = A.shape[0]
    rng = np.random.default_rng(4)
    b = rng.uniform(-1, 1, (n, 3))
    # factorization numbering: b_perm[new] = b[invp[new]]
    bp = b[invp]
    xp = orc.solve(S, 2 if lu else 0, L, U, bp)
    x = np.empty_like(xp)
    x[invp] = xp
    xr = spla.spsolve(sp.csc_matrix(A), b)
    assert _rel(x, xr) < 1e-11


def test_dense_matrix_flop_count_is_dense_cholesky():
    """The reference flop model (common/flops.h, symbol/symbol_cost_flops.c)
    summed over the cblks of a dense SPD matrix is the dense Cholesky count
    n^3/3 + n^2/2 + n/6 (LAPACK working note 41) plus the reference's TRSM
    quirk: FADDS_TRSM is defined as FMULS_TRMM (common/flops.h:74-75), so a
    one-column cblk with M rows below counts M extra flops, n(n-1)/2 in all
    for the personal ordering's one-column supernodes."""
    n = 40
    A = sp.csc_matrix(np.ones((n, n)) + n * np.eye(n))
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = pa.factotype.LLT
    ip[pa.iparm.verbose] = pa.verbose.Not
    ip[pa.iparm.ordering] = pa.ordering.Personal
    h = pa.init(ip, dp)
    try:
        spm = pa.spmatrix(A, mtxtype=pa.mtxtype.Symmetric)
        pa.task_analyze(h, spm, None)
        cblk, blok, perm, invp, dims = pa.get_structure(h)
        assert (cblk[:, 1] - cblk[:, 0] + 1).sum() == n
        th = float(dp[pa.dparm.fact_thflops])
    finally:
        pa.finalize(h)
    assert th == pytest.approx(n ** 3 / 3. + n ** 2 / 2. + n / 6. + n * (n - 1) / 2., rel=1e-12)


def read_rua(path):
    """Harwell-Boeing real unsymmetric assembled matrix (fixed-width fields
    given by the header's Fortran formats)."""
    with open(path) as f:
        lines = f.read().splitlines()
    totcrd, ptrcrd, indcrd, valcrd = [int(x) for x in lines[1].split()[:4]]
    nrow, ncol, nnz = [int(x) for x in lines[2][14:].split()[:3]]
    fmts = lines[3]
    body = lines[4:]
    ptr = " ".join(body[:ptrcrd]).split()
    ind = " ".join(body[ptrcrd:ptrcrd + indcrd]).split()
    vals = []
    w = int(fmts[fmts.index("E", 40) + 1:].split(".")[0]) if "E" in fmts[40:] else 16
    for l in body[ptrcrd + indcrd:ptrcrd + indcrd + valcrd]:
        vals += [float(l[i:i + w].replace("D", "E")) for i in range(0, len(l.rstrip()), w)]
    colptr = np.array(ptr, dtype=np.int64) - 1
    rows = np.array(ind, dtype=np.int64) - 1
    return sp.csc_matrix((np.array(vals[:nnz]), rows, colptr), shape=(nrow, ncol))


def test_orsirr_fixture_lu_is_doolittle():
    """The reference's LU test matrix (test/matrix/orsirr.rua, 1030 x 1030):
    the oracle's no-pivot LU on the analysed structure equals the dense
    Doolittle LU of the permuted matrix."""
    A = read_rua(os.path.join(HERE, "golden", "orsirr.rua"))
    assert A.shape == (1030, 1030) and A.nnz == 6858
    n = A.shape[0]
    # ordering: reverse Cuthill-McKee of the symmetrised pattern, given as a personal ordering
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    pat = (abs(A) + abs(A.T)).tocsr()
    rcm = reverse_cuthill_mckee(pat, symmetric_mode=True).astype(np.int64)
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = pa.factotype.LU
    ip[pa.iparm.verbose] = pa.verbose.Not
    h = pa.init(ip, dp)
    try:
        spm = pa.spmatrix(A, mtxtype=pa.mtxtype.General)
        pa.task_analyze(h, spm, pa.order.from_permutation(rcm))
        cblk, blok, perm, invp, dims = pa.get_structure(h)
    finally:
        pa.finalize(h)
    S = orc.Structure(cblk, blok)
    L, U, npiv = orc.factorize(S, A, perm, 2, 1e-300)
    Ap = A.toarray()[np.ix_(invp, invp)]
    Lr, Ur = doolittle(Ap)
    Ld, Ud = _expand(S, L, n), _expand(S, U, n).T
    assert _rel(np.tril(Ld, -1), np.tril(Lr, -1)) < 1e-10
    assert _rel(np.triu(Ud), Ur) < 1e-10
    assert npiv == 0
