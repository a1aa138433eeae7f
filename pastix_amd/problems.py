"""Synthetic sparse problems of the benchmark configurations.

Grid vertex (i, j, k) of an nx x ny x nz grid is numbered i + nx*j + nx*ny*k,
the numbering the grid nested dissection (``pastixOrderGrid``) assumes.
All matrices are returned as ``scipy.sparse.csc_matrix`` in FP64 with sorted
indices.  Random draws use ``numpy.random.default_rng(seed)``.
"""
import numpy as np
import scipy.sparse as sp


def _grid_index(nx, ny, nz):
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    return i.ravel(order="F"), j.ravel(order="F"), k.ravel(order="F")


def _stencil(nx, ny, nz, offsets, weights_fn):
    """Assemble sum over stencil offsets (dx, dy, dz) of weight * e_p e_q^T
    for neighbour q = p + offset inside the grid."""
    n = nx * ny * nz
    ii, jj, kk = _grid_index(nx, ny, nz)
    p = ii + nx * jj + nx * ny * kk
    rows, cols, vals = [], [], []
    for (dx, dy, dz) in offsets:
        i2, j2, k2 = ii + dx, jj + dy, kk + dz
        m = (i2 >= 0) & (i2 < nx) & (j2 >= 0) & (j2 < ny) & (k2 >= 0) & (k2 < nz)
        q = i2[m] + nx * j2[m] + nx * ny * k2[m]
        w = weights_fn(dx, dy, dz, p[m], q)
        rows.append(p[m])
        cols.append(q)
        vals.append(np.broadcast_to(w, q.shape).astype(np.float64))
    A = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return A


def laplacian(nx, ny=1, nz=1, shift=0.0):
    """(2*dim)-point Laplacian: diagonal 2*dim (+ shift), -1 to each grid
    neighbour, Dirichlet boundary (missing neighbours are dropped).  3D 7-point
    for nz > 1, 2D 5-point for nz == 1 < ny."""
    dim = 1 + (ny > 1) + (nz > 1)
    offs = [(0, 0, 0)]
    for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        if (d[1] != 0 and ny == 1) or (d[2] != 0 and nz == 1):
            continue
        offs.append(d)

    def w(dx, dy, dz, p, q):
        return (2.0 * dim + shift) if (dx, dy, dz) == (0, 0, 0) else -1.0
    return _stencil(nx, ny, nz, offs, w)


def convection_diffusion_27(nx, ny, nz, velocity=(0.3, 0.2, 0.1), diffusion=1.0):
    """27-point diffusion (centre 26*D, every one of the 26 neighbours -D) plus
    first-order upwind convection with a positive velocity v: v_d is added to
    the diagonal and -v_d to the upstream neighbour in direction d.  The
    pattern is symmetric, the values are not; the matrix is diagonally
    dominant (weakly inside, strictly at the boundary)."""
    vx, vy, vz = velocity
    offs = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]

    def w(dx, dy, dz, p, q):
        if (dx, dy, dz) == (0, 0, 0):
            return 26.0 * diffusion + vx + vy + vz
        v = -diffusion
        if (dx, dy, dz) == (-1, 0, 0):
            v -= vx
        elif (dx, dy, dz) == (0, -1, 0):
            v -= vy
        elif (dx, dy, dz) == (0, 0, -1):
            v -= vz
        return v
    return _stencil(nx, ny, nz, offs, w)


def anisotropic_conductivity(nx, ny, nz, seed=2):
    """Per-cell conductivity theta = exp(N(0,1)) drawn with the given seed."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.standard_normal(nx * ny * nz))


def anisotropic_diffusion(nx, ny, nz, theta, ratio_z=10.0):
    """7-point finite-volume diffusion -div(K grad u) with cell conductivity
    theta, K = diag(1, 1, ratio_z) theta.  The face coefficient between cells
    p and q is the arithmetic mean (theta_p + theta_q)/2 times the directional
    factor; a Dirichlet boundary face adds theta_p times the factor to the
    diagonal.  SPD.  The values are linear in theta: see
    ``anisotropic_diffusion_jacobian`` for the map theta -> A.data."""
    n = nx * ny * nz
    theta = np.asarray(theta, dtype=np.float64)
    ii, jj, kk = _grid_index(nx, ny, nz)
    p = ii + nx * jj + nx * ny * kk
    rows, cols, vals = [], [], []
    diag = np.zeros(n)
    for (dx, dy, dz), fac in (((1, 0, 0), 1.0), ((-1, 0, 0), 1.0), ((0, 1, 0), 1.0),
                              ((0, -1, 0), 1.0), ((0, 0, 1), ratio_z), ((0, 0, -1), ratio_z)):
        i2, j2, k2 = ii + dx, jj + dy, kk + dz
        m = (i2 >= 0) & (i2 < nx) & (j2 >= 0) & (j2 < ny) & (k2 >= 0) & (k2 < nz)
        q = i2[m] + nx * j2[m] + nx * ny * k2[m]
        c = fac * 0.5 * (theta[p[m]] + theta[q])
        rows.append(p[m]); cols.append(q); vals.append(-c)
        np.add.at(diag, p[m], c)
        np.add.at(diag, p[~m], fac * theta[p[~m]])
    rows.append(np.arange(n)); cols.append(np.arange(n)); vals.append(diag)
    A = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return A


def anisotropic_diffusion_jacobian(nx, ny, nz, ratio_z=10.0):
    """The sparse linear map of ``anisotropic_diffusion``: returns (pattern,
    J) with ``pattern`` the sorted CSC of A (values for theta = 1) and J an
    (nnz x n) CSR matrix such that A(theta).data = J @ theta, so that
    dLoss/dtheta = J^T dLoss/dA.data."""
    n = nx * ny * nz
    P = anisotropic_diffusion(nx, ny, nz, np.ones(n), ratio_z)
    keys = np.repeat(np.arange(n, dtype=np.int64), np.diff(P.indptr)) * n + P.indices.astype(np.int64)
    ii, jj, kk = _grid_index(nx, ny, nz)
    p = ii + nx * jj + nx * ny * kk
    er, ec, ev = [], [], []   # rows of J (CSC positions), columns (cells), coefficients

    def pos(r, c):
        return np.searchsorted(keys, c.astype(np.int64) * n + r)
    for (dx, dy, dz), fac in (((1, 0, 0), 1.0), ((-1, 0, 0), 1.0), ((0, 1, 0), 1.0),
                              ((0, -1, 0), 1.0), ((0, 0, 1), ratio_z), ((0, 0, -1), ratio_z)):
        i2, j2, k2 = ii + dx, jj + dy, kk + dz
        m = (i2 >= 0) & (i2 < nx) & (j2 >= 0) & (j2 < ny) & (k2 >= 0) & (k2 < nz)
        pm = p[m]
        q = i2[m] + nx * j2[m] + nx * ny * k2[m]
        off = pos(pm, q)
        dia = pos(pm, pm)
        for cell in (pm, q):
            er += [off, dia]
            ec += [cell, cell]
            ev += [np.full(len(pm), -0.5 * fac), np.full(len(pm), 0.5 * fac)]
        pb = p[~m]
        er.append(pos(pb, pb)); ec.append(pb); ev.append(np.full(len(pb), fac))
    J = sp.csr_matrix((np.concatenate(ev), (np.concatenate(er), np.concatenate(ec))), shape=(P.nnz, n))
    J.sum_duplicates()
    return P, J


def uniform(n, seed=0, nrhs=1, low=-1.0, high=1.0):
    """Vector (or n x nrhs column-major array) drawn from U[low, high]."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(low, high, size=(nrhs, n)).T
    return np.asfortranarray(x[:, 0] if nrhs == 1 else x)


def backward_error(A, x, b):
    """||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)."""
    r = b - A @ x
    nA = abs(A).sum(axis=1).max()
    return np.abs(r).max() / (nA * np.abs(x).max() + np.abs(b).max())
