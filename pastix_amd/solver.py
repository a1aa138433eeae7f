"""Python interface to the MI355X PaStiX engine.

Mirrors the reference's Python wrapper (reference
wrappers/python/examples/pypastix/solver.py and pastix.py): module-level
functions with the C API names (``initParam``, ``init``, ``task_analyze``,
``task_numfact``, ``task_solve``, ``task_refine``, ``finalize``) and a
``solver`` object with ``setup``/``solve``/``check``.  The matrix is given as
a scipy CSC matrix; everything numerical runs on the GPU through
``libpastix_amd.so``.
"""
import ctypes as C
import numpy as np
import scipy.sparse as sp

from . import _lib
from .enum import iparm, dparm, factotype as _ft, verbose as _verbose, trans as _trans, mtxtype as _mtx

PASTIX_SUCCESS = 0


class PastixError(RuntimeError):
    pass


def _check(rc, what):
    if rc != PASTIX_SUCCESS:
        raise PastixError("%s failed with error code %d" % (what, rc))


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class spmatrix(object):
    """Owns the arrays of an ``spmatrix_t`` built from a scipy sparse matrix
    (0-based CSC, FP64).  ``mtxtype`` follows SpM: 111 general,
    112 symmetric, 113 hermitian."""

    def __init__(self, A, mtxtype=None):
        A = sp.csc_matrix(A)
        A.sort_indices()
        n = A.shape[0]
        if A.shape[0] != A.shape[1]:
            raise ValueError("square matrix expected")
        self.colptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
        self.rowptr = np.ascontiguousarray(A.indices, dtype=np.int64)
        if A.dtype == np.float32:
            self.values = np.ascontiguousarray(A.data, dtype=np.float32)
            flt = 2
        else:
            self.values = np.ascontiguousarray(A.data, dtype=np.float64)
            flt = 3
        if mtxtype is None:
            mtxtype = _mtx.General
        s = _lib.spmatrix_t()
        s.mtxtype = int(mtxtype)
        s.flttype = flt
        s.fmttype = 0
        s.baseval = 0
        s.gN = s.n = s.gNexp = s.nexp = n
        s.gnnz = s.nnz = s.gnnzexp = s.nnzexp = len(self.values)
        s.dof = 1
        s.layout = 102
        s.colptr = self.colptr.ctypes.data
        s.rowptr = self.rowptr.ctypes.data
        s.values = self.values.ctypes.data
        s.clustnbr = 1
        s.replicated = 1
        self.spm = s
        self.n = n
        self.A = A

    @property
    def ref(self):
        return C.byref(self.spm)


class order(object):
    """A ``pastix_order_t`` owned by Python (filled by the C library)."""

    def __init__(self):
        self.o = _lib.pastix_order_t()
        self._ptr = C.pointer(self.o)

    @staticmethod
    def grid(nx, ny=1, nz=1):
        """Closed-form nested dissection of a regular grid (pastixOrderGrid)."""
        o = order()
        rc = _lib.load().pastixOrderGrid(C.byref(o._ptr), nx, ny, nz)
        _check(rc, "pastixOrderGrid")
        return o

    @staticmethod
    def from_permutation(perm):
        """A personal ordering given by perm[old] = new only; the analysis
        builds the supernode partition from the elimination tree."""
        perm = np.asarray(perm, dtype=np.int64)
        n = len(perm)
        o = order()
        lib = _lib.load()
        _check(lib.pastixOrderAlloc(C.byref(o.o), n, 0), "pastixOrderAlloc")
        pt = np.ctypeslib.as_array(o.o.permtab, (n,))
        it = np.ctypeslib.as_array(o.o.peritab, (n,))
        pt[:] = perm
        it[perm] = np.arange(n)
        return o

    @staticmethod
    def from_partition(perm, rangtab, treetab):
        """A personal ordering with its supernode partition and tree
        (perm[old] = new, rangtab[cblknbr + 1], treetab[cblknbr], -1 for a
        root): the analysis uses the partition as given."""
        perm = np.asarray(perm, dtype=np.int64)
        rangtab = np.asarray(rangtab, dtype=np.int64)
        treetab = np.asarray(treetab, dtype=np.int64)
        n, c = len(perm), len(rangtab) - 1
        o = order()
        lib = _lib.load()
        _check(lib.pastixOrderAlloc(C.byref(o.o), n, c), "pastixOrderAlloc")
        np.ctypeslib.as_array(o.o.permtab, (n,))[:] = perm
        np.ctypeslib.as_array(o.o.peritab, (n,))[perm] = np.arange(n)
        np.ctypeslib.as_array(o.o.rangtab, (c + 1,))[:] = rangtab
        np.ctypeslib.as_array(o.o.treetab, (c,))[:] = treetab
        return o

    def arrays(self):
        n, c = self.o.vertnbr, self.o.cblknbr
        return (np.ctypeslib.as_array(self.o.permtab, (n,)).copy(),
                np.ctypeslib.as_array(self.o.peritab, (n,)).copy(),
                np.ctypeslib.as_array(self.o.rangtab, (c + 1,)).copy(),
                np.ctypeslib.as_array(self.o.treetab, (c,)).copy())

    def __del__(self):
        try:
            _lib.load().pastixOrderExit(C.byref(self.o))
        except Exception:
            pass


def initParam():
    ip = np.zeros(_lib.IPARM_SIZE, dtype=np.int64)
    dp = np.zeros(_lib.DPARM_SIZE, dtype=np.float64)
    _lib.load().pastixInitParam(_ptr(ip), _ptr(dp))
    return ip, dp


def init(ip, dp):
    handle = C.c_void_p()
    _lib.load().pastixInit(C.byref(handle), 0, _ptr(ip), _ptr(dp))
    if not handle.value:
        raise PastixError("pastixInit failed")
    return handle


def finalize(handle):
    _lib.load().pastixFinalize(C.byref(handle))


def subtask_order(handle, spm, o=None):
    _check(_lib.load().pastix_subtask_order(handle, spm.ref, C.byref(o.o) if o is not None else None),
           "pastix_subtask_order")


def task_analyze(handle, spm, o=None):
    lib = _lib.load()
    if o is not None:
        subtask_order(handle, spm, o)
        _check(lib.pastix_subtask_symbfact(handle), "pastix_subtask_symbfact")
        _check(lib.pastix_subtask_reordering(handle), "pastix_subtask_reordering")
        _check(lib.pastix_subtask_blend(handle), "pastix_subtask_blend")
    else:
        _check(lib.pastix_task_analyze(handle, spm.ref), "pastix_task_analyze")


def task_numfact(handle, spm):
    _check(_lib.load().pastix_task_numfact(handle, spm.ref), "pastix_task_numfact")


def _rhs_dtype(spm):
    """Arithmetic of the right-hand sides: that of the matrix (FP32 for a
    single precision spm, FP64 otherwise), as the reference's solve reads it."""
    return np.float32 if spm.values.dtype == np.float32 else np.float64


def task_solve(handle, spm, x):
    """In-place solve of the host array x (n or n x nrhs, column major, in
    the matrix's arithmetic); returns x (a converted copy if x had another
    type or layout)."""
    x = np.asfortranarray(x, dtype=_rhs_dtype(spm))
    nrhs = 1 if x.ndim == 1 else x.shape[1]
    _check(_lib.load().pastix_task_solve(handle, spm.n, nrhs, _ptr(x), spm.n), "pastix_task_solve")
    return x


def task_refine(handle, spm, b, x):
    """Refine x (in place, same arithmetic and layout as task_solve) for b."""
    b = np.asfortranarray(b, dtype=_rhs_dtype(spm))
    if x.dtype != b.dtype or not x.flags.f_contiguous:
        raise ValueError("x must be a column-major array of the matrix's arithmetic")
    nrhs = 1 if b.ndim == 1 else b.shape[1]
    _check(_lib.load().pastix_task_refine(handle, spm.n, nrhs, _ptr(b), spm.n, _ptr(x), spm.n),
           "pastix_task_refine")
    return x


def get_schedule(handle):
    """(tptr, ttsk) of the analysis' static CPU schedule: thread t runs the
    cblks ttsk[tptr[t]:tptr[t+1]] in order."""
    lib = _lib.load()
    nt = np.zeros(1, dtype=np.int64)
    _check(lib.pastix_amd_get_schedule(handle, _ptr(nt), None, None), "pastix_amd_get_schedule")
    dims = np.zeros(8, dtype=np.int64)
    _check(lib.pastix_amd_get_dims(handle, _ptr(dims)), "pastix_amd_get_dims")
    tptr = np.zeros(int(nt[0]) + 1, dtype=np.int64)
    ttsk = np.zeros(int(dims[1]), dtype=np.int64)
    _check(lib.pastix_amd_get_schedule(handle, _ptr(nt), _ptr(tptr), _ptr(ttsk)), "pastix_amd_get_schedule")
    return tptr, ttsk


def get_structure(handle):
    """(cblk[cblknbr,5], blok[bloknbr,5], perm, invp, dims) of the analysed
    problem; cblk rows: fcolnum, lcolnum, first block, stride, coefficient
    offset; blok rows: frownum, lrownum, lcblknm, fcblknm, coefind."""
    lib = _lib.load()
    dims = np.zeros(8, dtype=np.int64)
    _check(lib.pastix_amd_get_dims(handle, _ptr(dims)), "pastix_amd_get_dims")
    n, cblknbr, bloknbr = int(dims[0]), int(dims[1]), int(dims[2])
    cblk = np.zeros((cblknbr, 5), dtype=np.int64)
    blok = np.zeros((bloknbr, 5), dtype=np.int64)
    perm = np.zeros(n, dtype=np.int64)
    invp = np.zeros(n, dtype=np.int64)
    _check(lib.pastix_amd_get_structure(handle, _ptr(cblk), _ptr(blok), _ptr(perm), _ptr(invp)),
           "pastix_amd_get_structure")
    return cblk, blok, perm, invp, dims


TUNE_NCAND = 8   # PASTIX_AMD_TUNE_NCAND
TUNE_CANDIDATES = [  # {tm_min, narrow, small, tall64}: the order of the library's candidate list
    (1 << 30, 1, 1, 0), (384, 1, 1, 0), (256, 1, 1, 0), (512, 1, 1, 0), (1 << 30, 0, 1, 0), (1 << 30, 1, 0, 0),
    (384, 1, 1, 1), (256, 1, 1, 1)]


def tune(handle):
    """Autotune the launch schedule on the matrix of the last numfact
    (``pastix_amd_tune``): returns {"best": index, "params": [tm_min,
    narrow, small, tall64], "times": seconds per candidate (0: not
    applicable)}.
    The handle keeps the chosen schedule and its factors."""
    lib = _lib.load()
    times = np.zeros(TUNE_NCAND)
    params = np.zeros(4, dtype=np.int64)
    best = lib.pastix_amd_tune(handle, _ptr(times), _ptr(params))
    if best < 0:
        raise PastixError("pastix_amd_tune failed with error code %d" % -best)
    return {"best": int(best), "params": [int(v) for v in params], "times": [float(t) for t in times]}


def set_tuning(handle, params):
    """Select a schedule (params as returned by ``tune``) for the next
    factorizations."""
    p = np.ascontiguousarray(list(params) + [0] * (4 - len(params)), dtype=np.int64)
    _check(_lib.load().pastix_amd_set_tuning(handle, _ptr(p)), "pastix_amd_set_tuning")


def get_coeftab(handle, side=0):
    lib = _lib.load()
    dims = np.zeros(8, dtype=np.int64)
    _check(lib.pastix_amd_get_dims(handle, _ptr(dims)), "pastix_amd_get_dims")
    out = np.zeros(int(dims[3]), dtype=np.float64)
    _check(lib.pastix_amd_get_coeftab(handle, side, _ptr(out)), "pastix_amd_get_coeftab")
    return out


class solver(object):
    """Object interface of the reference wrapper (pypastix ``solver``).

    kwargs: ``verbose`` (0/1/2), ``factotype`` (``factotype.LLT`` or ``LU``),
    ``mixed`` (FP32 factorization + FP64 refinement), ``refine``
    (``refine.GMRES``/``CG``/``SR``), ``grid`` ((nx, ny, nz) for the grid
    nested dissection; otherwise the elimination-tree ordering is used).
    """

    def __init__(self, A=None, **kwargs):
        self.iparm, self.dparm = initParam()
        v = kwargs.setdefault("verbose", _verbose.No)
        self.iparm[iparm.verbose] = {0: _verbose.Not, 2: _verbose.Yes}.get(int(v), _verbose.No)
        self.factotype = kwargs.setdefault("factotype", _ft.LU)
        self.iparm[iparm.factorization] = self.factotype
        if kwargs.get("mixed"):
            self.iparm[iparm.mixed] = 1
        if "refine" in kwargs:
            self.iparm[iparm.refinement] = kwargs["refine"]
        self.grid = kwargs.get("grid")
        self.pastix_data = init(self.iparm, self.dparm)
        self.spmA = None
        if A is not None:
            self.setup(A)

    def setup(self, A, grid=None):
        grid = grid if grid is not None else self.grid
        mt = _mtx.Symmetric if self.factotype in (_ft.LLT, _ft.LLH) else _mtx.General
        self.spmA = spmatrix(A, mtxtype=mt)
        self.order = order.grid(*grid) if grid is not None else None
        task_analyze(self.pastix_data, self.spmA, self.order)
        task_numfact(self.pastix_data, self.spmA)

    def solve(self, b, refine=True):
        x = np.array(b, dtype=np.float64, order="F", copy=True)
        task_solve(self.pastix_data, self.spmA, x)
        if refine:
            task_refine(self.pastix_data, self.spmA, b, x)
        return x

    def check(self, x, b):
        from .problems import backward_error
        return backward_error(self.spmA.A, x, b)

    def finalize(self):
        if self.pastix_data is not None and self.pastix_data.value:
            finalize(self.pastix_data)
        self.pastix_data = None

    def __del__(self):
        try:
            self.finalize()
        except Exception:
            pass
