"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of bench.py as the checker; never by the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        P, I, D = C.c_void_p, C.c_int64, C.c_double
        for pfx in ("d", "s"):
            getattr(_lib, "oracle_%sfill" % pfx).argtypes = [I, P, P, P, P, P, P, P, P]
            getattr(_lib, "oracle_%spotrf" % pfx).argtypes = [I, P, P, P, D, P, P]
            getattr(_lib, "oracle_%sgetrf" % pfx).argtypes = [I, P, P, P, P, D, P]
            getattr(_lib, "oracle_%ssytrf" % pfx).argtypes = [I, P, P, P, P, D, P]
            getattr(_lib, "oracle_%ssolve" % pfx).argtypes = [I, P, P, C.c_int, C.c_int, P, P, I, P, I]
            getattr(_lib, "oracle_%sgemmsp_one" % pfx).argtypes = [P, P, I, I, C.c_int, P, P, P]
            getattr(_lib, "oracle_%spotrf_diag" % pfx).argtypes = [I, P, D, P, P]
            getattr(_lib, "oracle_%sgetrf_diag" % pfx).argtypes = [I, P, P, D, P]
            getattr(_lib, "oracle_%ssytrf_diag" % pfx).argtypes = [I, P, D, P]
            getattr(_lib, "oracle_%strsm_cblk" % pfx).argtypes = [P, P, I, C.c_int, P, P]
            getattr(_lib, "oracle_%sset_blas" % pfx).argtypes = [P, P]
            getattr(_lib, "oracle_%spotrf_par" % pfx).argtypes = [I, P, P, P, D, P, P, C.c_int]
            getattr(_lib, "oracle_%sgetrf_par" % pfx).argtypes = [I, P, P, P, P, D, P, C.c_int]
            getattr(_lib, "oracle_%sfactor_static" % pfx).argtypes = [I, P, P, P, P, C.c_int, D, P, P, C.c_int,
                                                                       P, P]
    return _lib


def blas_pointers(dtype=np.float64):
    """(gemm, trsm) Fortran-convention function pointers of the BLAS that
    SciPy exports to compiled code (scipy.linalg.cython_blas capsules)."""
    from scipy.linalg import cython_blas
    api = C.pythonapi
    api.PyCapsule_GetName.restype = C.c_char_p
    api.PyCapsule_GetName.argtypes = [C.py_object]
    api.PyCapsule_GetPointer.restype = C.c_void_p
    api.PyCapsule_GetPointer.argtypes = [C.py_object, C.c_char_p]
    pfx = "d" if dtype == np.float64 else "s"
    out = []
    for name in (pfx + "gemm", pfx + "trsm"):
        cap = cython_blas.__pyx_capi__[name]
        out.append(api.PyCapsule_GetPointer(cap, api.PyCapsule_GetName(cap)))
    return tuple(out)


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class Structure(object):
    """Supernodal structure as exported by pastix_amd_get_structure."""

    def __init__(self, cblk, blok):
        self.cblk = np.ascontiguousarray(cblk, dtype=np.int64)
        # sentinel row: first block of cblk k+1
        sent = np.zeros((1, 5), dtype=np.int64)
        sent[0, 2] = len(blok)
        self.cblkx = np.ascontiguousarray(np.vstack([self.cblk, sent]))
        self.blok = np.ascontiguousarray(blok, dtype=np.int64)
        self.cblknbr = len(cblk)
        last = self.cblk[-1]
        self.coefnbr = int(last[4] + last[3] * (last[1] - last[0] + 1)) if len(cblk) else 0


def permuted_csc(A, perm):
    """P A P^T as CSC (rows sorted) with the transpose values in the same
    pattern: (colptr, rowtab, Lv, Uv); perm[old] = new."""
    import scipy.sparse as sp
    A = sp.csc_matrix(A)
    pat = (abs(A) + abs(A.T)).tocsc()
    pat.data[:] = 1.0
    n = A.shape[0]
    Pm = sp.csc_matrix((np.ones(n), (perm, np.arange(n))), shape=(n, n))
    B = (Pm @ A @ Pm.T).tocsc()
    Bt = (Pm @ A.T @ Pm.T).tocsc()
    patp = (Pm @ pat @ Pm.T).tocsc()
    patp.sort_indices()
    rows = patp.indices.astype(np.int64)
    cols = np.repeat(np.arange(n), np.diff(patp.indptr))
    Lv = np.asarray(B[rows, cols]).ravel()
    Uv = np.asarray(Bt[rows, cols]).ravel()
    return patp.indptr.astype(np.int64), rows, Lv.astype(np.float64), Uv.astype(np.float64)


def factorize(S, A, perm, factotype, crit, dtype=np.float64):
    """Fill + sequential factorization; returns (L, U or None, npiv).
    factotype 0/3: LL^t; 1/4: LDL^t (U is the L D copy of the panels); 2: LU."""
    lib = load()
    pfx = "d" if dtype == np.float64 else "s"
    colptr, rowtab, Lv, Uv = permuted_csc(A, perm)
    L = np.zeros(S.coefnbr, dtype=dtype)
    U = np.zeros(S.coefnbr, dtype=dtype) if factotype in (1, 2, 4) else None
    getattr(lib, "oracle_%sfill" % pfx)(S.cblknbr, _p(S.cblkx), _p(S.blok), _p(colptr), _p(rowtab),
                                        _p(Lv), _p(Uv), _p(L), _p(U if factotype == 2 else None))
    npiv = np.zeros(1, dtype=np.int64)
    nneg = np.zeros(1, dtype=np.int64)
    if factotype == 2:
        getattr(lib, "oracle_%sgetrf" % pfx)(S.cblknbr, _p(S.cblkx), _p(S.blok), _p(L), _p(U), crit, _p(npiv))
    elif factotype in (1, 4):
        getattr(lib, "oracle_%ssytrf" % pfx)(S.cblknbr, _p(S.cblkx), _p(S.blok), _p(L), _p(U), crit, _p(npiv))
    else:
        getattr(lib, "oracle_%spotrf" % pfx)(S.cblknbr, _p(S.cblkx), _p(S.blok), _p(L), crit, _p(npiv), _p(nneg))
    return L, U, int(npiv[0])


def factorize_parallel(S, A, perm, crit, nthreads, use_blas=True, dtype=np.float64, factotype=0):
    """Fill + level-parallel LL^t (factotype 0) or LU (2) on nthreads
    threads, the dense kernels on SciPy's BLAS (one BLAS thread per task)
    when use_blas; returns (L, U or None, npiv, seconds of the factorization
    alone)."""
    import time
    lib = load()
    pfx = "d" if dtype == np.float64 else "s"
    colptr, rowtab, Lv, Uv = permuted_csc(A, perm)
    L = np.zeros(S.coefnbr, dtype=dtype)
    U = np.zeros(S.coefnbr, dtype=dtype) if factotype == 2 else None
    getattr(lib, "oracle_%sfill" % pfx)(S.cblknbr, _p(S.cblkx), _p(S.blok), _p(colptr), _p(rowtab),
                                        _p(Lv), _p(Uv), _p(L), _p(U))
    npiv = np.zeros(1, dtype=np.int64)
    nneg = np.zeros(1, dtype=np.int64)
    setb = getattr(lib, "oracle_%sset_blas" % pfx)
    try:
        if use_blas:
            g, t = blas_pointers(dtype)
            setb(C.c_void_p(g), C.c_void_p(t))
        try:
            from threadpoolctl import threadpool_limits
            limit = threadpool_limits(limits=1, user_api="blas")
        except Exception:
            limit = None
        t0 = time.perf_counter()
        if factotype == 2:
            getattr(lib, "oracle_%sgetrf_par" % pfx)(S.cblknbr, _p(S.cblkx), _p(S.blok), _p(L), _p(U), crit,
                                                     _p(npiv), int(nthreads))
        else:
            getattr(lib, "oracle_%spotrf_par" % pfx)(S.cblknbr, _p(S.cblkx), _p(S.blok), _p(L), crit, _p(npiv),
                                                     _p(nneg), int(nthreads))
        sec = time.perf_counter() - t0
        if limit is not None:
            limit.restore_original_limits()
    finally:
        setb(None, None)
    return L, U, int(npiv[0]), sec


def factorize_static(S, A, perm, crit, tptr, ttsk, use_blas=True, dtype=np.float64, factotype=0):
    """Fill + LL^t (factotype 0) or LU (2) by the threads of a static
    schedule: thread t runs the cblks ttsk[tptr[t]:tptr[t+1]] in order, each
    waiting for its incoming updates (the dense kernels on SciPy's BLAS, one
    BLAS thread per task, when use_blas); returns (L, U or None, npiv,
    seconds of the factorization alone)."""
    import time
    lib = load()
    pfx = "d" if dtype == np.float64 else "s"
    tptr = np.ascontiguousarray(tptr, dtype=np.int64)
    ttsk = np.ascontiguousarray(ttsk, dtype=np.int64)
    colptr, rowtab, Lv, Uv = permuted_csc(A, perm)
    L = np.zeros(S.coefnbr, dtype=dtype)
    U = np.zeros(S.coefnbr, dtype=dtype) if factotype == 2 else None
    getattr(lib, "oracle_%sfill" % pfx)(S.cblknbr, _p(S.cblkx), _p(S.blok), _p(colptr), _p(rowtab),
                                        _p(Lv), _p(Uv), _p(L), _p(U))
    npiv = np.zeros(1, dtype=np.int64)
    nneg = np.zeros(1, dtype=np.int64)
    setb = getattr(lib, "oracle_%sset_blas" % pfx)
    try:
        if use_blas:
            g, t = blas_pointers(dtype)
            setb(C.c_void_p(g), C.c_void_p(t))
        try:
            from threadpoolctl import threadpool_limits
            limit = threadpool_limits(limits=1, user_api="blas")
        except Exception:
            limit = None
        t0 = time.perf_counter()
        rc = getattr(lib, "oracle_%sfactor_static" % pfx)(S.cblknbr, _p(S.cblkx), _p(S.blok), _p(L), _p(U),
                                                          1 if factotype == 2 else 0, crit, _p(npiv), _p(nneg),
                                                          len(tptr) - 1, _p(tptr), _p(ttsk))
        sec = time.perf_counter() - t0
        if limit is not None:
            limit.restore_original_limits()
    finally:
        setb(None, None)
    if rc != 0:
        raise ValueError("the static schedule does not hold every cblk exactly once")
    return L, U, int(npiv[0]), sec


def solve(S, factotype, L, U, b_perm, trans=0):
    """Solve in the factorization numbering; b_perm is n or n x nrhs."""
    lib = load()
    dtype = L.dtype
    pfx = "d" if dtype == np.float64 else "s"
    x = np.array(b_perm, dtype=dtype, order="F", copy=True)
    nrhs = 1 if x.ndim == 1 else x.shape[1]
    n = x.shape[0]
    getattr(lib, "oracle_%ssolve" % pfx)(S.cblknbr, _p(S.cblkx), _p(S.blok), int(factotype), int(trans),
                                         _p(L), _p(U if U is not None else L), nrhs, _p(x), n)
    return x


def expand_L(S, L, n, side_unit=False):
    """Dense lower-triangular factor from the 2D tiled coeftab (small n)."""
    M = np.zeros((n, n), dtype=np.float64)
    for k in range(S.cblknbr):
        f, l, fb, stride, off = S.cblk[k]
        nc = l - f + 1
        lb = S.cblkx[k + 1, 2]
        for b in range(fb, lb):
            fr, lr, _, _, coef = S.blok[b]
            m = lr - fr + 1
            tile = L[off + coef: off + coef + m * nc].reshape((nc, m)).T
            M[fr:lr + 1, f:l + 1] = tile
    return np.tril(M)
