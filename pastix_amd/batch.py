"""Many independent sparse systems factored and solved in one device pass.

The systems are stacked into one block-diagonal matrix whose ordering is the
union of the per-system orderings (each system's grid nested dissection,
shifted by the system's offset; the supernode trees become one forest).  The
level schedule of the engine then groups the supernodes of the same height
of every system into the same diagonal / TRSM / GEMM launches, so the whole
batch is factored with one launch sequence of forest-height length and solved
with one multi-RHS sweep.
"""
import numpy as np
import scipy.sparse as sp

from . import _lib
from .enum import factotype as _ft, mtxtype as _mtx, verbose as _verbose, iparm as _ip
from .solver import (_check, init, finalize, initParam, spmatrix, task_analyze, task_numfact, task_solve,
                     order)


def forest_order(orders, sizes):
    """Union of per-system orderings (``order`` objects with supernode
    partitions) into one ``order`` for the block-diagonal matrix."""
    lib = _lib.load()
    parts = [o.arrays() for o in orders]
    n = int(sum(sizes))
    ncb = int(sum(len(p[3]) for p in parts))
    out = order()
    _check(lib.pastixOrderAlloc(C_byref(out.o), n, ncb), "pastixOrderAlloc")
    perm = np.ctypeslib.as_array(out.o.permtab, (n,))
    invp = np.ctypeslib.as_array(out.o.peritab, (n,))
    rang = np.ctypeslib.as_array(out.o.rangtab, (ncb + 1,))
    tree = np.ctypeslib.as_array(out.o.treetab, (ncb + 1,))
    off, coff = 0, 0
    for (p, ip, r, t), m in zip(parts, sizes):
        c = len(t)
        perm[off:off + m] = p + off
        invp[off:off + m] = ip + off
        rang[coff:coff + c] = r[:c] + off
        tree[coff:coff + c] = np.where(t >= 0, t + coff, -1)
        off += m
        coff += c
    rang[ncb] = n
    return out


def C_byref(x):
    import ctypes
    return ctypes.byref(x)


class BatchedSolver(object):
    """Factor (LL^t, FP64) and solve a list of SPD systems in one pass.

    ``mats``: scipy sparse SPD matrices; ``grids``: per matrix the (nx, ny,
    nz) grid of its unknowns (for its nested dissection)."""

    def __init__(self, mats, grids, mixed=False):
        self.sizes = [m.shape[0] for m in mats]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.A = sp.block_diag([sp.csc_matrix(m) for m in mats], format="csc")
        self.A.sort_indices()
        self.iparm, self.dparm = initParam()
        self.iparm[_ip.factorization] = _ft.LLT
        self.iparm[_ip.verbose] = _verbose.Not
        if mixed:
            self.iparm[_ip.mixed] = 1
        self.h = init(self.iparm, self.dparm)
        self.spm = spmatrix(self.A, mtxtype=_mtx.Symmetric)
        self._orders = [order.grid(*g) for g in grids]
        self.order = forest_order(self._orders, self.sizes)
        task_analyze(self.h, self.spm, self.order)

    def factorize(self):
        task_numfact(self.h, self.spm)

    def solve(self, rhs):
        """rhs: list of (n_i x nrhs) arrays (same nrhs); returns the list of
        solutions.  One solve call for the stacked right-hand sides."""
        B = np.asfortranarray(np.vstack(rhs))
        task_solve(self.h, self.spm, B)
        return [B[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self.sizes))]

    def close(self):
        if self.h is not None:
            finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def uneven_laplacians(count=256, low=64, high=512, seed=1):
    """The batch of the batching workload: 2D 5-point Laplacians on n_i x n_i
    grids with n_i ~ U{low..high} plus a diagonal shift ~ U[0,1] (seed)."""
    from .problems import laplacian
    rng = np.random.default_rng(seed)
    ns = rng.integers(low, high + 1, size=count)
    shifts = rng.uniform(0., 1., size=count)
    mats = [laplacian(int(m), int(m), 1, shift=float(s)) for m, s in zip(ns, shifts)]
    grids = [(int(m), int(m), 1) for m in ns]
    return mats, grids
