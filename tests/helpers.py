"""Shared set-up of the parity tests: analyse a problem with the product,
factor it on the GPU and with the CPU oracle on the same structure."""
import numpy as np
import scipy.sparse as sp

import pastix_amd as pa
from oracle import pyoracle as orc


class Problem(object):
    def __init__(self, A, grid, factotype, mixed=False, blocksize=None, analyze_only=False, order=None):
        self.A = sp.csc_matrix(A)
        self.n = A.shape[0]
        self.ft = factotype
        self.lu = (factotype == pa.factotype.LU)
        self.iparm, self.dparm = pa.initParam()
        self.iparm[pa.iparm.factorization] = factotype
        self.iparm[pa.iparm.verbose] = pa.verbose.Not
        if mixed:
            self.iparm[pa.iparm.mixed] = 1
        if blocksize is not None:
            self.iparm[pa.iparm.min_blocksize] = blocksize[0]
            self.iparm[pa.iparm.max_blocksize] = blocksize[1]
            self.iparm[pa.iparm.tasks2d_width] = blocksize[0]
        self.h = pa.init(self.iparm, self.dparm)
        mt = pa.mtxtype.General if self.lu else pa.mtxtype.Symmetric
        self.spm = pa.spmatrix(self.A, mtxtype=mt)
        self.order = pa.order.grid(*grid) if grid is not None else order
        pa.task_analyze(self.h, self.spm, self.order)
        self.cblk, self.blok, self.perm, self.invp, self.dims = pa.get_structure(self.h)
        self.S = orc.Structure(self.cblk, self.blok)
        self.crit = np.finfo(np.float64).eps / 2 * sp.linalg.norm(self.A)
        if not analyze_only:
            pa.task_numfact(self.h, self.spm)

    def oracle_factors(self, dtype=np.float64):
        return orc.factorize(self.S, self.A, self.perm, 2 if self.lu else 0, self.crit, dtype=dtype)

    def gpu_factors(self):
        L = pa.get_coeftab(self.h, 0)
        U = pa.get_coeftab(self.h, 1) if self.lu else None
        return L, U

    def lower_mask(self):
        """Coefficients of the coeftab that hold factor entries: all of the
        off-diagonal tiles and the lower triangle of the diagonal tiles (the
        upper triangle of an LL^t diagonal tile is never read)."""
        mask = np.ones(self.S.coefnbr, dtype=bool)
        for k in range(self.S.cblknbr):
            f, l, fb, stride, off = self.cblk[k]
            nc = l - f + 1
            tile = mask[off:off + nc * nc].reshape((nc, nc)).T  # [row, col]
            tile[np.triu_indices(nc, 1)] = False
            mask[off:off + nc * nc] = tile.T.ravel()
        return mask

    def close(self):
        pa.finalize(self.h)
