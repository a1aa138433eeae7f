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
    """A version-0 file (partition and permutation, no tree): as in the
    reference (order/pastix_subtask_order.c:362-381 runs when the tree is
    missing), the analysis rebuilds the partition from the permutation
    (fundamental supernodes + amalgamation), exactly as for the same
    permutation given as a personal ordering."""
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
    assert s2[4][0] == n
    perm = np.array(body[cblknbr + 1:], dtype=np.int64)
    ip, dp = pa.initParam()
    ip[pa.iparm.factorization] = pa.factotype.LLT
    ip[pa.iparm.verbose] = pa.verbose.Not
    h = pa.init(ip, dp)
    try:
        spm = pa.spmatrix(A, mtxtype=pa.mtxtype.Symmetric)
        pa.task_analyze(h, spm, pa.order.from_permutation(perm))
        s3 = pa.get_structure(h)
        th3 = float(dp[pa.dparm.fact_thflops])
    finally:
        pa.finalize(h)
    for a, b in zip(s2[:4], s3[:4]):
        assert np.array_equal(a, b)
    assert th2 == th3


# ---------------------------------------------------------------------------
# Supernodes of a bare permutation (reference order/order_find_supernodes.c,
# order/order_amalgamate.c, symbol/fax_csr_amalgamate.c), called through the
# library's internal entry points pa_order_find_supernodes /
# pa_order_amalgamate and checked against the textbook definitions computed
# by brute force.
# ---------------------------------------------------------------------------
import ctypes as C

from pastix_amd import _lib


def _sym_pattern(A):
    pat = (abs(A) + abs(A.T)).tocsc()
    pat.setdiag(0)
    pat.eliminate_zeros()
    pat.sort_indices()
    return pat


def _supernodes(pat, perm0, lvl=None):
    """(perm, rangtab, treetab, rangtab before amalgamation) from the
    library for the permutation perm0 of the symmetric pattern."""
    lib = _lib.load()
    n = pat.shape[0]
    o = _lib.pastix_order_t()
    assert lib.pastixOrderAlloc(C.byref(o), n, 0) == 0
    np.ctypeslib.as_array(o.permtab, (n,))[:] = perm0
    np.ctypeslib.as_array(o.peritab, (n,))[perm0] = np.arange(n)
    cp = np.ascontiguousarray(pat.indptr, dtype=np.int64)
    rw = np.ascontiguousarray(pat.indices, dtype=np.int64)
    f = lib.pa_order_find_supernodes
    f.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]
    assert f(C.byref(o), n, cp.ctypes.data, rw.ctypes.data) == 0
    r0 = np.ctypeslib.as_array(o.rangtab, (o.cblknbr + 1,)).copy()
    p0 = np.ctypeslib.as_array(o.permtab, (n,)).copy()
    if lvl is not None:
        g = lib.pa_order_amalgamate
        g.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int, C.c_int64, C.c_int64]
        assert g(C.byref(o), n, cp.ctypes.data, rw.ctypes.data, 0, lvl[0], lvl[1]) == 0
    c = o.cblknbr
    out = (np.ctypeslib.as_array(o.permtab, (n,)).copy(), np.ctypeslib.as_array(o.rangtab, (c + 1,)).copy(),
           np.ctypeslib.as_array(o.treetab, (c,)).copy(), r0, p0)
    assert lib.pastixOrderCheck(C.byref(o)) == 0
    lib.pastixOrderExit(C.byref(o))
    return out


def _textbook(pat, perm):
    """Elimination tree parent[] and the row sets struct[j] (rows > j) of the
    Cholesky factor of P pat P^T, by symbolic elimination."""
    n = pat.shape[0]
    invp = np.empty(n, dtype=np.int64)
    invp[perm] = np.arange(n)
    B = pat[invp][:, invp].tocsc()
    B.sort_indices()
    parent = -np.ones(n, dtype=np.int64)
    struct = [None] * n
    kids = [[] for _ in range(n)]
    for j in range(n):
        s = set(int(r) for r in B.indices[B.indptr[j]:B.indptr[j + 1]] if r > j)
        for c in kids[j]:
            s |= {r for r in struct[c] if r > j}
        struct[j] = s
        if s:
            parent[j] = min(s)
            kids[parent[j]].append(j)
    return parent, struct


def _random_spd_pattern(n, density, seed):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=rng, format="csc")
    return _sym_pattern(M + M.T + sp.eye(n)), rng.permutation(n).astype(np.int64)


@pytest.mark.parametrize("case", ["grid2d", "random", "orsirr"])
def test_fundamental_supernodes_of_a_bare_permutation(case):
    """The permutation is replaced by a postorder of the elimination tree of
    the permuted matrix (same tree, siblings kept in their elimination
    order); every supernode start is a textbook fundamental-supernode start
    (a vertex without exactly one child, or a leaf of a row subtree); every
    textbook start is found except inside the subtree of column 0, where the
    reference's previous-row markers start at column 0
    (order/order_find_supernodes.c:437-455); the tree links each supernode to
    the one holding the etree parent of its last column."""
    if case == "grid2d":
        A = problems.laplacian(9, 7, 1)
        pat = _sym_pattern(A)
        perm0 = np.random.default_rng(0).permutation(A.shape[0]).astype(np.int64)
    elif case == "random":
        pat, perm0 = _random_spd_pattern(300, 0.01, 1)
    else:
        from test_cpu_oracle import read_rua
        from scipy.sparse.csgraph import reverse_cuthill_mckee
        pat = _sym_pattern(read_rua(os.path.join(os.path.dirname(__file__), "golden", "orsirr.rua")))
        rcm = reverse_cuthill_mckee(pat.tocsr(), symmetric_mode=True)
        perm0 = np.empty(pat.shape[0], dtype=np.int64)
        perm0[rcm] = np.arange(pat.shape[0])
    n = pat.shape[0]
    perm, rang, tree, r0, p0 = _supernodes(pat, perm0)
    assert np.array_equal(rang, r0)
    parent, struct = _textbook(pat, perm)
    parent0, _ = _textbook(pat, perm0)
    # same tree on the vertices, relabelled by a postorder
    inv0 = np.empty(n, dtype=np.int64)
    inv0[perm0] = np.arange(n)
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    for v in range(n):
        a, b = parent0[perm0[v]], parent[perm[v]]
        assert (a < 0 and b < 0) or (a >= 0 and b >= 0 and inv0[a] == inv[b])
    size = np.ones(n, dtype=np.int64)
    first = np.arange(n)
    for j in range(n):
        if parent[j] >= 0:
            assert parent[j] > j
            size[parent[j]] += size[j]
            first[parent[j]] = min(first[parent[j]], first[j])
    assert np.array_equal(first, np.arange(n) - size + 1)    # subtrees are contiguous
    # siblings in their previous elimination order
    for j in range(n):
        kids = [c for c in range(first[j], j) if parent[c] == j]
        assert list(np.argsort(perm0[inv[kids]])) == list(range(len(kids)))
    nch = np.zeros(n, dtype=np.int64)
    for j in range(n):
        if parent[j] >= 0:
            nch[parent[j]] += 1
    textbook = {0} | {j for j in range(1, n)
                      if not (parent[j - 1] == j and nch[j] == 1 and struct[j - 1] - {j} == struct[j])}
    starts = set(int(x) for x in rang[:-1])
    assert starts <= textbook
    missed = textbook - starts
    assert all(first[j] == 0 and nch[j] == 1 for j in missed), sorted(missed)[:10]
    col2sn = np.repeat(np.arange(len(rang) - 1), np.diff(rang))
    for k in range(len(rang) - 1):
        p = parent[rang[k + 1] - 1]
        assert tree[k] == (col2sn[p] if p >= 0 else -1)


def _partition_nnz(pat, perm, rang):
    """nnz(L) of the supernodal factor of a partition: each supernode holds
    the union of its columns' row sets on all of its columns."""
    parent, struct = _textbook(pat, perm)
    nnz = 0
    for k in range(len(rang) - 1):
        a, b = int(rang[k]), int(rang[k + 1])
        rows = set()
        for j in range(a, b):
            rows |= {r for r in struct[j] if r >= b}
        nc = b - a
        nnz += nc * (nc + 1) // 2 + nc * len(rows)
    return nnz


@pytest.mark.parametrize("lvl", [(0, 0), (5, 5), (10, 10), (2, -8)], ids=["0", "5", "10", "2-blas8"])
@pytest.mark.parametrize("case", ["random", "orsirr"])
def test_amalgamation_merges_whole_supernodes_within_the_fill_budget(case, lvl):
    """Amalgamation (reference symbol/fax_csr_amalgamate.c): the new
    partition is a union of whole supernodes along the supernode tree, its
    elimination order is still valid (every etree parent after its child),
    and with a zero budget only merges that add no explicit zero happen.  The
    budget of each phase (lround(nnz(L) * lvl_cblk / 100) structural, then
    lround(nnz(L) * |lvl_blas| / 100) driven by the BLAS time model; a
    negative lvl_blas caps the structural phase at 2 %) is charged with the
    fill computed when a merge is queued; the reference refreshes that cost
    for the father and the merged node's sons only
    (symbol/fax_csr_amalgamate.c:680-716), not for its siblings, so the zeros
    actually added may exceed the budget somewhat: they are checked against
    twice the budget."""
    if case == "random":
        pat, perm0 = _random_spd_pattern(300, 0.01, 1)
    else:
        from test_cpu_oracle import read_rua
        from scipy.sparse.csgraph import reverse_cuthill_mckee
        pat = _sym_pattern(read_rua(os.path.join(os.path.dirname(__file__), "golden", "orsirr.rua")))
        rcm = reverse_cuthill_mckee(pat.tocsr(), symmetric_mode=True)
        perm0 = np.empty(pat.shape[0], dtype=np.int64)
        perm0[rcm] = np.arange(pat.shape[0])
    n = pat.shape[0]
    perm, rang, tree, r0, p0 = _supernodes(pat, perm0, lvl)
    nnz0 = _partition_nnz(pat, p0, r0)
    nnz1 = _partition_nnz(pat, perm, rang)
    c, b = lvl
    cb = (min(abs(b), 2) if b < 0 else c) / 100.
    budget = int(round(nnz0 * cb)) + (int(round(nnz0 * abs(b) / 100.)) if abs(b) / 100. > cb else 0)
    assert 0 <= nnz1 - nnz0 <= 2 * budget
    if lvl == (0, 0):
        assert nnz1 == nnz0
    # whole old supernodes in each new one
    inv0 = np.empty(n, dtype=np.int64)
    inv0[p0] = np.arange(n)
    new_of = np.repeat(np.arange(len(rang) - 1), np.diff(rang))[perm[inv0]]   # by old column
    for k in range(len(r0) - 1):
        assert len(set(new_of[r0[k]:r0[k + 1]])) == 1
    assert len(rang) - 1 <= len(r0) - 1
    parent, _ = _textbook(pat, perm)
    assert all(parent[j] < 0 or parent[j] > j for j in range(n))
    assert all(tree[k] == -1 or tree[k] > k for k in range(len(tree)))
