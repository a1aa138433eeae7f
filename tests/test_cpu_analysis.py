"""CPU checks of the host analysis that drives the device: the grid nested
dissection (order/order_grids.c restated), the supernodal symbolic
factorization and the SolverMatrix layout (blend/solver_matrix_gen*.c).
Pinned by invariants that a wrong index or a dropped block would break:
the separator property of every dissection node, containment of the exact
Cholesky fill pattern, and the 2D tile offsets."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import pastix_amd as pa
from pastix_amd import problems

GRIDS = [(3, 3, 3), (5, 4, 3), (7, 7, 7), (10, 9, 8)]


def _analyse(A, grid, blocksize=None, factotype=None):
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = pa.factotype.LLT if factotype is None else factotype
    ip[pa.iparm.verbose] = pa.verbose.Not
    if blocksize:
        ip[pa.iparm.min_blocksize], ip[pa.iparm.max_blocksize] = blocksize
        ip[pa.iparm.tasks2d_width] = blocksize[0]
    h = pa.init(ip, dp)
    try:
        spm = pa.spmatrix(A, mtxtype=pa.mtxtype.Symmetric)
        pa.task_analyze(h, spm, pa.order.grid(*grid))
        return pa.get_structure(h)
    finally:
        pa.finalize(h)


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_grid_nested_dissection_separates(grid):
    """Every supernode of the dissection separates the subtrees of its
    children: no edge of the grid graph joins two different child subtrees."""
    o = pa.order.grid(*grid)
    perm, invp, rang, tree = o.arrays()
    n = int(np.prod(grid))
    assert np.array_equal(perm[invp], np.arange(n))
    assert np.all(np.diff(rang) > 0) and rang[0] == 0 and rang[-1] == n
    ncb = len(tree)
    assert all(tree[c] == -1 or tree[c] > c for c in range(ncb))
    # supernode of each new index, then the subtree root below each node
    sn = np.repeat(np.arange(ncb), np.diff(rang))
    children = [[] for _ in range(ncb)]
    for c in range(ncb):
        if tree[c] >= 0:
            children[tree[c]].append(c)
    A = problems.laplacian(*grid).tocoo()
    u, v = perm[A.row], perm[A.col]           # edges in the new numbering
    su, sv = sn[u], sn[v]

    def ancestors(c):
        out = []
        while c != -1:
            out.append(c)
            c = tree[c]
        return out
    anc = [set(ancestors(c)) for c in range(ncb)]
    assert sum(len(ch) >= 2 for ch in children) > 0     # a real dissection tree
    for p in range(ncb):
        if len(children[p]) < 2:
            continue
        # which child subtree (if any) each supernode belongs to
        lab = {}
        for ci in children[p]:
            for c in range(ncb):
                if ci in anc[c]:
                    lab[c] = ci
        for a, b in zip(su, sv):
            if a in lab and b in lab:
                assert lab[a] == lab[b], (p, a, b)


def _dense_fill(A, invp):
    """Nonzero pattern of the exact Cholesky factor of P A P^T (random SPD
    values on A's pattern, so no cancellation)."""
    n = A.shape[0]
    rng = np.random.default_rng(0)
    B = sp.triu(A, 1).tocoo()
    vals = rng.uniform(0.1, 1.0, B.nnz)
    S = sp.coo_matrix((vals, (B.row, B.col)), shape=(n, n))
    S = S + S.T
    D = S.toarray() + np.diag(np.abs(S).sum(axis=1).A.ravel() + 1.)
    L = np.linalg.cholesky(D[np.ix_(invp, invp)])
    return np.abs(L) > 1e-14 * np.abs(L).max()


@pytest.mark.parametrize("grid", GRIDS[:3], ids=lambda g: "x".join(map(str, g)))
@pytest.mark.parametrize("blocksize", [None, (4, 8)], ids=["default", "split"])
def test_symbolic_structure_covers_the_cholesky_fill(grid, blocksize):
    A = problems.laplacian(*grid)
    cblk, blok, perm, invp, dims = _analyse(A, grid, blocksize)
    n = A.shape[0]
    covered = np.zeros((n, n), dtype=bool)
    nb = len(blok)
    for k in range(len(cblk)):
        f, l, fb, stride, off = cblk[k]
        lb = cblk[k + 1, 2] if k + 1 < len(cblk) else nb
        for b in range(fb, lb):
            covered[blok[b, 0]:blok[b, 1] + 1, f:l + 1] = True
    fill = _dense_fill(A, invp)
    assert not np.any(fill & ~covered), "fill entries outside the symbolic structure"


@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
@pytest.mark.parametrize("blocksize", [None, (4, 8)], ids=["default", "split"])
def test_solver_matrix_layout(grid, blocksize):
    A = problems.laplacian(*grid)
    cblk, blok, perm, invp, dims = _analyse(A, grid, blocksize)
    n = A.shape[0]
    nb = len(blok)
    ncb = len(cblk)
    assert dims[0] == n and dims[1] == ncb and dims[2] == nb
    # the cblks partition the columns in order
    assert cblk[0, 0] == 0 and cblk[-1, 1] == n - 1
    assert np.all(cblk[1:, 0] == cblk[:-1, 1] + 1)
    widths = cblk[:, 1] - cblk[:, 0] + 1
    if blocksize:
        assert widths.max() <= blocksize[1]
    coef = 0
    col2cblk = np.repeat(np.arange(ncb), widths)
    for k in range(ncb):
        f, l, fb, stride, off = cblk[k]
        nc = l - f + 1
        lb = cblk[k + 1, 2] if k + 1 < ncb else nb
        assert off == coef                              # panels are contiguous
        coef += stride * nc
        # first block is the diagonal block
        assert blok[fb, 0] == f and blok[fb, 1] == l and blok[fb, 2] == k and blok[fb, 3] == k
        rowoff = 0
        prev = l
        for b in range(fb, lb):
            fr, lr, lc, fc, ci = blok[b]
            assert lc == k and fr <= lr
            assert ci == rowoff * nc                    # 2D tile: coefind = rows above * nc
            rowoff += lr - fr + 1
            if b > fb:
                assert fr > prev                        # sorted, disjoint
                prev = lr
                # the rows of an off-diagonal block lie inside its facing cblk
                assert fc > k and col2cblk[fr] == fc and col2cblk[lr] == fc
        assert rowoff == stride
    assert dims[3] == coef


def test_lu_analysis_uses_the_symmetrised_pattern():
    """A nonsymmetric matrix is analysed on the pattern of A + A^T: the
    structure is the same as for the symmetric Laplacian of that pattern."""
    grid = (5, 4, 3)
    C = problems.convection_diffusion_27(*grid)
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = pa.factotype.LU
    ip[pa.iparm.verbose] = pa.verbose.Not
    h = pa.init(ip, dp)
    try:
        spm = pa.spmatrix(C, mtxtype=pa.mtxtype.General)
        pa.task_analyze(h, spm, pa.order.grid(*grid))
        s1 = pa.get_structure(h)
    finally:
        pa.finalize(h)
    pat = (abs(C) + abs(C.T)).tocsc()
    pat.data[:] = 1.
    S = sp.csc_matrix(pat + sp.eye(C.shape[0]) * 30.)
    s2 = _analyse(S, grid, factotype=pa.factotype.LU)
    for a, b in zip(s1[:4], s2[:4]):
        assert np.array_equal(a, b)


def _structure(A, grid, io, path, lu=False):
    old = os.environ.get("PASTIX_FILE_ORDER")
    os.environ["PASTIX_FILE_ORDER"] = path
    try:
        ip, dp = pa.initParam()
        ip[pa.iparm.factorization] = pa.factotype.LU if lu else pa.factotype.LLT
        ip[pa.iparm.verbose] = pa.verbose.Not
        ip[pa.iparm.io_strategy] = io
        h = pa.init(ip, dp)
        try:
            spm = pa.spmatrix(A, mtxtype=pa.mtxtype.General if lu else pa.mtxtype.Symmetric)
            pa.task_analyze(h, spm, pa.order.grid(*grid) if grid else None)
            return pa.get_structure(h), float(dp[pa.dparm.fact_thflops])
        finally:
            pa.finalize(h)
    finally:
        if old is None:
            del os.environ["PASTIX_FILE_ORDER"]
        else:
            os.environ["PASTIX_FILE_ORDER"] = old


def test_ordering_file_round_trip(tmp_path):
    """An ordering saved with IPARM_IO_STRATEGY = Save (the reference's text
    format, order/order_io.c) and loaded back with Load gives the same
    supernodal structure, without the grid ordering being supplied."""
    grid = (9, 8, 7)
    A = problems.laplacian(*grid)
    path = str(tmp_path / "ordergen")
    (s1, th1) = _structure(A, grid, pa.io.Save, path)
    txt = open(path).read().split()
    assert txt[0] == "1" and int(txt[2]) == A.shape[0]
    (s2, th2) = _structure(A, None, pa.io.Load, path)
    for a, b in zip(s1[:4], s2[:4]):
        assert np.array_equal(a, b)
    assert th1 == th2


def test_ordering_file_version0_without_tree(tmp_path):
    """A version-0 file (partition and permutation, no tree) is completed
    with the supernode tree of the elimination tree, as the reference's
    loader leaves it to the analysis: the structure matches the one of the
    same partition with its tree."""
    grid = (8, 7, 6)
    A = problems.laplacian(*grid)
    path = str(tmp_path / "ordergen")
    (s1, th1) = _structure(A, grid, pa.io.Save, path)
    tok = open(path).read().split()
    cblknbr, n = int(tok[1]), int(tok[2])
    body = tok[3:3 + (cblknbr + 1) + n]          # rangtab + permtab, no treetab
    p0 = str(tmp_path / "v0")
    with open(p0, "w") as f:
        f.write("0\n%d\t%d\n%s\n" % (cblknbr, n, "\n".join(body)))
    (s2, th2) = _structure(A, None, pa.io.Load, p0)
    assert s2[4][0] == n and len(s2[0]) == len(s1[0])
    assert th1 == th2
    for a, b in zip(s1[:4], s2[:4]):
        assert np.array_equal(a, b)
