"""MI355X-native PaStiX factor+solve engine.

The numerical path (coeftab fill, supernodal LL^t / LU factorization with
static pivoting, triangular solves, SpMV and FP64 iterative refinement) runs
in hand-written HIP kernels for gfx950 inside ``lib/libpastix_amd.so``; the
integer analysis (grid nested dissection, symbolic factorization, splitting,
solver structure) runs on the host in C.  This package holds the ctypes
binding, the reference-style Python wrapper and the differentiable solve.
"""
from . import enum
from .enum import iparm, dparm, factotype, refine, verbose, trans, mtxtype, ordering, io, dir, error, side, uplo, diag
from .solver import (solver, spmatrix, order, initParam, init, finalize, task_analyze,
                     task_numfact, task_solve, task_refine, get_structure, get_schedule, get_coeftab,
                     tune, set_tuning, PastixError)
from . import problems

__all__ = ["enum", "iparm", "dparm", "factotype", "refine", "verbose", "trans", "mtxtype", "ordering", "io",
           "dir", "error", "side", "uplo", "diag",
           "solver", "spmatrix", "order", "initParam", "init", "finalize", "task_analyze",
           "task_numfact", "task_solve", "task_refine", "get_structure", "get_schedule", "get_coeftab",
           "tune", "set_tuning", "PastixError", "problems"]
