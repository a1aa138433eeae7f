"""Differentiable sparse solve x = A^{-1} b on the MI355X engine.

``SparseSolver`` holds one analysis (ordering, symbolic factorization and
device structure) for a fixed CSC pattern; every forward call refactors with
the current values and solves on HBM-resident tensors.  After the first
factorization (which builds the permuted CSC on the host), new values go
from the device tensor straight into the device CSC
(``pastix_amd_values_dev``), so a forward pass moves no matrix data over
PCIe.  The backward pass reuses the factors:

    lambda   = A^{-T} g          (transposed device solve, refined in FP64
                                 with the transposed operator)
    dL/db    = lambda
    dL/dA_ij = -lambda_i x_j     only on the stored pattern (HIP kernel
                                 ``pastix_amd_sampled_outer``)

Tensors are FP64 on the HIP device; values are in the order of the CSC
pattern given at construction.
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp
import torch

from . import _lib
from .enum import factotype as _ft, mtxtype as _mtx, verbose as _verbose, iparm as _ip, trans as _tr
from .solver import PastixError, _check, init, finalize, initParam, spmatrix, task_analyze, task_numfact, order


def _dptr(t):
    return C.c_void_p(int(t.data_ptr()))


class SparseSolver(object):
    """One analysed pattern; ``symmetric`` selects LL^t (SPD) or LU."""

    def __init__(self, pattern, symmetric=True, grid=None, refine=True):
        P = sp.csc_matrix(pattern)
        P.sort_indices()
        self.n = P.shape[0]
        self.indptr = np.ascontiguousarray(P.indptr, dtype=np.int64)
        self.indices = np.ascontiguousarray(P.indices, dtype=np.int64)
        self.nnz = int(self.indptr[-1])
        self.symmetric = symmetric
        self.refine = refine
        self.iparm, self.dparm = initParam()
        self.iparm[_ip.factorization] = _ft.LLT if symmetric else _ft.LU
        self.iparm[_ip.verbose] = _verbose.Not
        self.h = init(self.iparm, self.dparm)
        self._mt = _mtx.Symmetric if symmetric else _mtx.General
        A0 = sp.csc_matrix((np.ones(self.nnz), self.indices, self.indptr), shape=P.shape)
        spm = spmatrix(A0, mtxtype=self._mt)
        task_analyze(self.h, spm, order.grid(*grid) if grid is not None else None)
        dev = torch.device("cuda")
        self.d_colptr = torch.from_numpy(self.indptr).to(dev)
        self.d_rows = torch.from_numpy(self.indices).to(dev)
        self._spm = None
        self.version = 0   # bumped by every factorization

    def factor(self, values):
        """Numerical factorization with `values` (nnz, CSC order)."""
        self.version += 1
        if values.shape[0] != self.nnz:
            raise ValueError("expected %d values, got %d" % (self.nnz, values.shape[0]))
        if self._spm is None or not values.is_cuda:
            v = values.detach().to("cpu", torch.float64).numpy()
            A = sp.csc_matrix((v, self.indices, self.indptr), shape=(self.n, self.n))
            self._spm = spmatrix(A, mtxtype=self._mt)   # kept alive: the value map is built from it
            task_numfact(self.h, self._spm)
            return
        lib = _lib.load()
        v = values.detach().to(torch.float64).contiguous()
        torch.cuda.current_stream(v.device).synchronize()
        _check(lib.pastix_amd_values_dev(self.h, _dptr(v)), "pastix_amd_values_dev")
        _check(lib.pastix_subtask_bcsc2ctab(self.h), "pastix_subtask_bcsc2ctab")
        _check(lib.pastix_subtask_sopalin(self.h), "pastix_subtask_sopalin")

    def solve(self, b, transpose=False):
        """x = op(A)^{-1} b for a device FP64 tensor b (n or n x nrhs)."""
        lib = _lib.load()
        b = b.detach().to(torch.float64)
        nrhs = 1 if b.dim() == 1 else b.shape[1]
        bcm = b.contiguous() if b.dim() == 1 else b.t().contiguous()   # column-major copy
        x = bcm.clone()
        torch.cuda.current_stream(x.device).synchronize()
        tr = _tr.Trans if transpose else _tr.NoTrans
        _check(lib.pastix_amd_solve_dev(self.h, tr, nrhs, _dptr(x), self.n), "pastix_amd_solve_dev")
        # FP64 refinement of op(A) x = b with the same operator
        if self.refine:
            it = lib.pastix_amd_refine_dev_adv(self.h, tr, nrhs, _dptr(bcm), self.n, _dptr(x), self.n)
            if it < 0:
                raise PastixError("pastix_amd_refine_dev_adv failed with error code %d" % it)
        return x if b.dim() == 1 else x.t().contiguous()

    def sampled_outer(self, u, v, alpha=-1.0):
        """alpha * sum_r u[row(e), r] v[col(e), r] on the pattern."""
        lib = _lib.load()
        uu = u.t().contiguous() if u.dim() == 2 else u.contiguous()
        vv = v.t().contiguous() if v.dim() == 2 else v.contiguous()
        nrhs = 1 if u.dim() == 1 else u.shape[1]
        out = torch.empty(self.nnz, dtype=torch.float64, device=u.device)
        torch.cuda.current_stream(out.device).synchronize()
        _check(lib.pastix_amd_sampled_outer(self.h, _dptr(self.d_colptr), _dptr(self.d_rows), nrhs, _dptr(uu),
                                            self.n, _dptr(vv), self.n, alpha, _dptr(out)),
               "pastix_amd_sampled_outer")
        return out

    def close(self):
        if self.h is not None:
            finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, solver):
        solver.factor(values)
        x = solver.solve(b)
        ctx.solver = solver
        ctx.version = solver.version
        ctx.save_for_backward(x, values.detach())
        return x

    @staticmethod
    def backward(ctx, g):
        x, values = ctx.saved_tensors
        s = ctx.solver
        if s.version != ctx.version:
            # the solver was refactored since this forward: the adjoint
            # needs the factors of the forward's values
            s.factor(values)
            ctx.version = s.version
        lam = s.solve(g.contiguous(), transpose=True)
        gv = s.sampled_outer(lam, x, -1.0) if ctx.needs_input_grad[0] else None
        gb = lam if ctx.needs_input_grad[1] else None
        return gv, gb, None


def solve(solver, values, b):
    """x = A(values)^{-1} b, differentiable in values (CSC order) and b."""
    return _Solve.apply(values, b, solver)
