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
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
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


def _dense_thflops(n, o):
    A = sp.csc_matrix(np.ones((n, n)) + n * np.eye(n))
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = pa.factotype.LLT
    ip[pa.iparm.verbose] = pa.verbose.Not
    ip[pa.iparm.ordering] = pa.ordering.Personal
    h = pa.init(ip, dp)
    try:
        spm = pa.spmatrix(A, mtxtype=pa.mtxtype.Symmetric)
        pa.task_analyze(h, spm, o)
        cblk, blok, perm, invp, dims = pa.get_structure(h)
        assert (cblk[:, 1] - cblk[:, 0] + 1).sum() == n
        return float(dp[pa.dparm.fact_thflops]), len(cblk)
    finally:
        pa.finalize(h)


def test_dense_matrix_flop_count_is_dense_cholesky():
    """The reference flop model (common/flops.h, symbol/symbol_cost_flops.c)
    summed over the cblks of a dense SPD matrix is the dense Cholesky count
    n^3/3 + n^2/2 + n/6 (LAPACK working note 41).  Without a partition the
    analysis finds one supernode (fundamental supernodes of a dense matrix);
    with a partition of one-column supernodes the reference's TRSM quirk
    shows: FADDS_TRSM is defined as FMULS_TRMM (common/flops.h:74-75), so a
    one-column cblk with M rows below counts M extra flops, n(n-1)/2 in all."""
    n = 40
    dense = n ** 3 / 3. + n ** 2 / 2. + n / 6.
    th, ncblk = _dense_thflops(n, None)
    assert ncblk == 1
    assert th == pytest.approx(dense, rel=1e-12)
    tree = np.append(np.arange(1, n), -1)
    th1, ncblk1 = _dense_thflops(n, pa.order.from_partition(np.arange(n), np.arange(n + 1), tree))
    assert ncblk1 == n
    assert th1 == pytest.approx(dense + n * (n - 1) / 2., rel=1e-12)


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


# ---------------------------------------------------------------------------
# FP32 restatement (oracle_s*): the same source compiled with REAL=float.
# Pinned by float32 LAPACK / a float32 Doolittle and by the pinned FP64
# functions on the same inputs, with tolerances of a few n*eps32: an FP64
# computation rounded at the end, a cast to a narrower type, or another pivot
# rule fails them.
# ---------------------------------------------------------------------------
EPS32 = float(np.finfo(np.float32).eps)


def doolittle32(A):
    """No-pivot LU in float32 arithmetic, column by column."""
    n = A.shape[0]
    U = np.array(A, dtype=np.float32)
    L = np.eye(n, dtype=np.float32)
    for k in range(n):
        L[k + 1:, k] = U[k + 1:, k] / U[k, k]
        U[k + 1:, k:] -= np.outer(L[k + 1:, k], U[k, k:]).astype(np.float32)
        U[k + 1:, k] = 0.
    return L, U


@pytest.mark.parametrize("nc", [1, 5, 63, 64, 65, 130])
def test_spotrf_diag_is_float_lapack_cholesky(nc):
    from scipy.linalg import lapack
    rng = np.random.default_rng(nc)
    M = rng.uniform(-1, 1, (nc, nc))
    M = (M @ M.T + nc * np.eye(nc)).astype(np.float32)
    D = np.asfortranarray(np.tril(M)).ravel(order="F").copy()
    npiv = np.zeros(1, dtype=np.int64)
    nneg = np.zeros(1, dtype=np.int64)
    orc.load().oracle_spotrf_diag(nc, orc._p(D), 1e-30, orc._p(npiv), orc._p(nneg))
    L = np.tril(D.reshape((nc, nc), order="F"))
    Ls, info = lapack.spotrf(M, lower=1)
    assert info == 0
    L64 = np.linalg.cholesky(M.astype(np.float64))
    tol = 4 * nc * EPS32
    assert _rel(L, np.tril(Ls)) < tol
    # single precision arithmetic: within the FP32 bound of the exact factor
    assert _rel(L, L64) < tol
    if nc >= 63:
        assert _rel(L, L64) > 1e-12   # not an FP64 computation
    assert npiv[0] == 0 and nneg[0] == 0


def test_spotrf_static_pivot_closed_form():
    """Float static pivot: diag(1e-20, 4) with crit 1e-6 gives
    diag(sqrt(1e-6), 2) and one replaced pivot."""
    D = np.array([1e-20, 0., 0., 4.], dtype=np.float32)
    npiv = np.zeros(1, dtype=np.int64)
    nneg = np.zeros(1, dtype=np.int64)
    orc.load().oracle_spotrf_diag(2, orc._p(D), 1e-6, orc._p(npiv), orc._p(nneg))
    assert np.isclose(D[0], np.sqrt(1e-6), rtol=4 * EPS32) and D[3] == 2.
    assert npiv[0] == 1


@pytest.mark.parametrize("nc", [1, 7, 64, 65, 100])
def test_sgetrf_diag_is_float_doolittle(nc):
    rng = np.random.default_rng(100 + nc)
    M = (rng.uniform(-1, 1, (nc, nc)) + nc * np.eye(nc)).astype(np.float32)
    Lt = np.asfortranarray(np.tril(M)).ravel(order="F").copy()
    Ut = np.asfortranarray(np.tril(M.T, -1)).ravel(order="F").copy()
    npiv = np.zeros(1, dtype=np.int64)
    orc.load().oracle_sgetrf_diag(nc, orc._p(Lt), orc._p(Ut), 1e-30, orc._p(npiv))
    F = Lt.reshape((nc, nc), order="F")
    Lr, Ur = doolittle32(M)
    tol = 8 * nc * EPS32
    assert _rel(np.tril(F, -1), np.tril(Lr, -1)) < tol
    assert _rel(np.triu(F), Ur) < tol
    assert _rel(Ut.reshape((nc, nc), order="F"), F.T) == 0.
    assert npiv[0] == 0


def test_sgetrf_signed_static_pivot():
    """The LU static pivot keeps the sign (reference
    kernels/core_zgetrfsp.c:81-91): a -1e-20 pivot with crit 1e-3 becomes
    -1e-3 in both precisions, and both count one pivot."""
    for pfx, dt in (("d", np.float64), ("s", np.float32)):
        M = np.array([[-1e-20, 1.], [2., 3.]], dtype=dt)
        Lt = np.asfortranarray(np.tril(M)).ravel(order="F").copy()
        Ut = np.asfortranarray(np.tril(M.T, -1)).ravel(order="F").copy()
        npiv = np.zeros(1, dtype=np.int64)
        getattr(orc.load(), "oracle_%sgetrf_diag" % pfx)(2, orc._p(Lt), orc._p(Ut), 1e-3, orc._p(npiv))
        F = Lt.reshape((2, 2), order="F")
        assert npiv[0] == 1
        assert F[0, 0] == dt(-1e-3)
        assert np.isclose(F[1, 0], 2. / -1e-3, rtol=4 * EPS32)
        assert np.isclose(F[1, 1], 3. - (2. / -1e-3) * 1., rtol=4 * EPS32)


@pytest.mark.parametrize("lu", [False, True], ids=["llt", "lu"])
def test_single_factorization_and_solve_follow_the_double_ones(lu):
    """oracle_s{potrf,getrf,solve} on the analysed structure against the
    pinned oracle_d*: the factors agree to a few n*eps32 relative (the
    Laplacian and the diagonally dominant convection-diffusion matrix have
    bounded growth), a solve with the FP32 factors has a backward error at
    the FP32 level, and the same pivots are replaced.  A crit that fires on
    the FP32 side only would show as a pivot count mismatch."""
    grid = (12, 11, 10)
    A = problems.convection_diffusion_27(*grid) if lu else problems.laplacian(*grid)
    S, perm, invp, th = _analyse(A, grid, lu=lu)
    n = A.shape[0]
    crit = np.finfo(np.float64).eps / 2 * sp.linalg.norm(A)
    ft = 2 if lu else 0
    Ld, Ud, npd = orc.factorize(S, A, perm, ft, crit)
    Ls, Us, nps = orc.factorize(S, A, perm, ft, crit, dtype=np.float32)
    assert nps == npd == 0
    tol = 50 * EPS32 * np.sqrt(n)
    if lu:
        assert _rel(_expand(S, Ls.astype(np.float64), n), _expand(S, Ld, n)) < tol
        assert _rel(_expand(S, Us.astype(np.float64), n), _expand(S, Ud, n)) < tol
    else:
        assert _rel(np.tril(_expand(S, Ls.astype(np.float64), n)), np.tril(_expand(S, Ld, n))) < tol
        # genuinely single precision
        assert _rel(np.tril(_expand(S, Ls.astype(np.float64), n)), np.tril(_expand(S, Ld, n))) > 1e-10
    rng = np.random.default_rng(6)
    b = rng.uniform(-1, 1, n)
    for trans in ((0, 1) if lu else (0,)):
        xs = orc.solve(S, ft, Ls, Us, b[invp].astype(np.float32), trans=trans).astype(np.float64)
        xd = orc.solve(S, ft, Ld, Ud, b[invp], trans=trans)
        x = np.empty(n)
        x[invp] = xs
        Aop = A.T if trans else A
        berr = problems.backward_error(sp.csc_matrix(Aop), x, b)
        assert berr < 20 * EPS32, berr
        assert _rel(xs, xd) < 1e3 * EPS32


def test_single_panel_kernels_follow_the_double_ones():
    """oracle_strsm_cblk and oracle_sgemmsp_one against the pinned FP64
    kernels on the same float-representable inputs."""
    grid = (9, 8, 7)
    A = problems.laplacian(*grid)
    S, perm, invp, th = _analyse(A, grid, lu=False)
    lib = orc.load()
    rng = np.random.default_rng(9)
    # the widest cblk with blocks below its diagonal
    k = max((k for k in range(S.cblknbr) if S.cblkx[k + 1, 2] - S.cblkx[k, 2] >= 3),
            key=lambda k: S.cblk[k, 1] - S.cblk[k, 0])
    f, l, fb, stride, off = [int(v) for v in S.cblk[k]]
    nc = l - f + 1
    Dm = rng.uniform(-1, 1, (nc, nc)) + 4 * nc * np.eye(nc)
    D32 = np.asfortranarray(Dm).ravel(order="F").astype(np.float32)
    panel32 = rng.uniform(-1, 1, stride * nc).astype(np.float32)
    for mode in (0, 1, 2):
        ps, pd_ = panel32.copy(), panel32.astype(np.float64)
        lib.oracle_strsm_cblk(orc._p(S.cblkx), orc._p(S.blok), k, mode, orc._p(D32), orc._p(ps))
        D64 = D32.astype(np.float64)
        lib.oracle_dtrsm_cblk(orc._p(S.cblkx), orc._p(S.blok), k, mode, orc._p(D64), orc._p(pd_))
        assert _rel(ps.astype(np.float64), pd_) < 20 * nc * EPS32, mode
    L32 = rng.uniform(-1, 1, S.coefnbr).astype(np.float32)
    for j in range(fb + 1, int(S.cblkx[k + 1, 2])):
        Cs, Cd = L32.copy(), L32.astype(np.float64)
        lib.oracle_sgemmsp_one(orc._p(S.cblkx), orc._p(S.blok), k, j, 0, orc._p(L32), orc._p(L32), orc._p(Cs))
        L64 = L32.astype(np.float64)
        lib.oracle_dgemmsp_one(orc._p(S.cblkx), orc._p(S.blok), k, j, 0, orc._p(L64), orc._p(L64), orc._p(Cd))
        assert _rel(Cs.astype(np.float64), Cd) < 4 * nc * EPS32
        assert np.abs(Cd - L64).max() > 0   # the update touched something


def test_transposed_oracle_solve_matches_sparse_direct_solve():
    """oracle_dsolve(trans=1) on an LU factorization solves A^T x = b
    (U^T y = b, then L^T x = y; reference sopalin/pastix_task_solve.c:543-557)."""
    grid = (6, 5, 4)
    A = problems.convection_diffusion_27(*grid)
    S, perm, invp, th = _analyse(A, grid, lu=True)
    L, U, _ = orc.factorize(S, A, perm, 2, 1e-300)
    n = A.shape[0]
    rng = np.random.default_rng(11)
    b = rng.uniform(-1, 1, (n, 2))
    xp = orc.solve(S, 2, L, U, b[invp], trans=1)
    x = np.empty_like(xp)
    x[invp] = xp
    xr = spla.spsolve(sp.csc_matrix(A.T), b)
    assert _rel(x, xr) < 1e-11
    # and it is not the non-transposed solve
    assert _rel(x, spla.spsolve(sp.csc_matrix(A), b)) > 1e-3


# ---------------------------------------------------------------------------
# LDL^t (oracle_*sytrf): pinned by the unique no-pivot LDL^t of the permuted
# matrix, read off a Doolittle LU (U = D L^t for a symmetric matrix).
# ---------------------------------------------------------------------------
def _ldlt_dense(M):
    L, U = doolittle(M)
    return L, np.diag(U).copy()


@pytest.mark.parametrize("nc", [1, 5, 63, 64, 65, 130])
def test_sytrf_diag_is_the_unique_ldlt(nc):
    """oracle_dsytrf_diag (reference kernels/core_zsytrfsp.c:68-215) on a
    symmetric indefinite block: strictly lower part L, diagonal D, strictly
    upper part D L^t."""
    rng = np.random.default_rng(200 + nc)
    M = rng.uniform(-1, 1, (nc, nc))
    M = M + M.T + np.diag(np.where(np.arange(nc) % 2, 1., -1.) * nc)   # symmetric, indefinite
    D = np.asfortranarray(np.tril(M)).ravel(order="F").copy()
    npiv = np.zeros(1, dtype=np.int64)
    orc.load().oracle_dsytrf_diag(nc, orc._p(D), 1e-300, orc._p(npiv))
    F = D.reshape((nc, nc), order="F")
    Lr, dr = _ldlt_dense(M)
    assert _rel(np.tril(F, -1), np.tril(Lr, -1)) < 1e-12
    assert _rel(np.diag(F), dr) < 1e-12
    assert _rel(np.triu(F, 1), np.triu((dr[:, None] * Lr.T), 1)) < 1e-12
    assert npiv[0] == 0
    if nc > 1:
        assert (np.diag(F) < 0).any() and (np.diag(F) > 0).any()


def test_sytrf_static_pivot_keeps_the_sign():
    """A pivot below the criterion becomes +-criterion (reference
    kernels/core_zsytrfsp.c:81-90): diag(-1e-20, 4) with crit 1e-10 gives
    D = (-1e-10, 4) and one replaced pivot."""
    D = np.array([-1e-20, 0., 0., 4.])
    npiv = np.zeros(1, dtype=np.int64)
    orc.load().oracle_dsytrf_diag(2, orc._p(D), 1e-10, orc._p(npiv))
    assert D[0] == -1e-10 and D[3] == 4. and npiv[0] == 1


def test_ldlt_factorization_is_the_unique_ldlt_and_solves():
    """Whole LDL^t of a symmetric indefinite matrix (the Laplacian shifted
    across part of its spectrum) on the analysed structure: L and D equal the
    dense no-pivot LDL^t of P A P^T, the panel copy equals L D, and the
    L / D / L^t solve matches a sparse direct solve."""
    grid = (6, 5, 4)
    A = (problems.laplacian(*grid) - 2.5 * sp.eye(120)).tocsc()
    S, perm, invp, th = _analyse(A, grid, lu=False)
    L, DLt, npiv = orc.factorize(S, A, perm, 1, 1e-300)
    assert npiv == 0
    n = A.shape[0]
    Ap = A.toarray()[np.ix_(invp, invp)]
    Lr, dr = _ldlt_dense(Ap)
    assert (dr < 0).any() and (dr > 0).any()
    F = _expand(S, L, n)
    assert _rel(np.tril(F, -1), np.tril(Lr, -1)) < 1e-11
    assert _rel(np.diag(F), dr) < 1e-11
    # panels of the copy: L_panel * D (columns scaled by the factor's own D)
    dF = np.diag(F)
    for k in range(S.cblknbr):
        f, l, fb, stride, off = S.cblk[k]
        nc = l - f + 1
        for b in range(fb + 1, S.cblkx[k + 1, 2]):
            fr, lr, _, _, coef = S.blok[b]
            m = lr - fr + 1
            X = L[off + coef: off + coef + m * nc].reshape((nc, m)).T
            Y = DLt[off + coef: off + coef + m * nc].reshape((nc, m)).T
            assert np.array_equal(Y, X * dF[f:l + 1][None, :])
    rng = np.random.default_rng(12)
    b = rng.uniform(-1, 1, (n, 2))
    xp = orc.solve(S, 1, L, DLt, b[invp])
    x = np.empty_like(xp)
    x[invp] = xp
    assert _rel(x, spla.spsolve(A, b)) < 1e-11
