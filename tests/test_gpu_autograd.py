"""Differentiable solve (pastix_amd.autograd): adjoint gradients against
central finite differences, the adjoint solve against the CPU oracle, and
the sampled outer product kernel against a dense product."""
import numpy as np
import pytest
import scipy.sparse as sp

import pastix_amd as pa
from pastix_amd import problems
from oracle import pyoracle as orc
from helpers import Problem

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def test_sampled_outer_product_on_the_pattern():
    torch = _torch()
    from pastix_amd.autograd import SparseSolver
    grid = (7, 6, 5)
    A = problems.laplacian(*grid)
    s = SparseSolver(A, symmetric=True, grid=grid)
    try:
        n = A.shape[0]
        rng = np.random.default_rng(11)
        u = rng.uniform(-1, 1, (n, 3))
        v = rng.uniform(-1, 1, (n, 3))
        got = s.sampled_outer(torch.from_numpy(u).cuda(), torch.from_numpy(v).cuda(), -2.0).cpu().numpy()
        C = sp.csc_matrix(A)
        C.sort_indices()
        cols = np.repeat(np.arange(n), np.diff(C.indptr))
        ref = -2.0 * np.einsum("ij,ij->i", u[C.indices], v[cols])
        assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    finally:
        s.close()


def _fd_product(xp, xm, xt, h):
    """(L(+h) - L(-h)) / 2h for L = 0.5 ||x - xt||^2, written as a product of
    differences so that the cancellation happens in x, not in L."""
    return float(np.dot(xp - xm, xp + xm - 2. * xt)) / (4. * h)


def test_adjoint_gradient_anisotropic_diffusion_64():
    """Anisotropic diffusion on 64^3 (n = 262144), theta = exp(N(0,1)) seed 2,
    ratio 10 in z, LL^t; loss 0.5 ||x - x_target||^2 with b ~ U[-1,1] (seed 3)
    and x_target ~ U[-1,1] (seed 4).  dL/dtheta on 32 random cells (seed 5)
    and dL/db on 8 random rows against central differences with h = 1e-6."""
    torch = _torch()
    from pastix_amd.autograd import SparseSolver, solve
    nx = ny = nz = 64
    n = nx * ny * nz
    theta = problems.anisotropic_conductivity(nx, ny, nz, seed=2)
    P, J = problems.anisotropic_diffusion_jacobian(nx, ny, nz, ratio_z=10.0)
    b = problems.uniform(n, seed=3)
    xt = problems.uniform(n, seed=4)
    s = SparseSolver(P, symmetric=True, grid=(nx, ny, nz))
    try:
        vals = torch.from_numpy(J @ theta).cuda().requires_grad_(True)
        bt = torch.from_numpy(b).cuda().requires_grad_(True)
        xtt = torch.from_numpy(xt).cuda()
        x = solve(s, vals, bt)
        loss = 0.5 * ((x - xtt) ** 2).sum()
        loss.backward()
        g_vals = vals.grad.cpu().numpy()
        g_theta = J.T @ g_vals
        g_b = bt.grad.cpu().numpy()
        x0 = x.detach().cpu().numpy()
        # the adjoint: A^T lambda = x - x_target
        A = sp.csc_matrix((J @ theta, P.indices, P.indptr), shape=(n, n))
        r = A.T @ g_b - (x0 - xt)
        assert np.abs(r).max() <= 1e-10 * np.abs(x0 - xt).max()

        h = 1e-6
        rng = np.random.default_rng(5)
        cells = rng.choice(n, 32, replace=False)
        fd = []
        for i in cells:
            sol = []
            for sgn in (1., -1.):
                th = theta.copy()
                th[i] += sgn * h
                s.factor(torch.from_numpy(J @ th))
                sol.append(s.solve(torch.from_numpy(b).cuda()).cpu().numpy())
            fd.append(_fd_product(sol[0], sol[1], xt, h))
        fd = np.array(fd)
        an = g_theta[cells]
        scale = np.abs(an).max()
        assert np.abs(fd - an).max() <= 1e-4 * scale, (np.abs(fd - an).max(), scale)

        s.factor(torch.from_numpy(J @ theta))
        rows = rng.choice(n, 8, replace=False)
        for i in rows:
            sol = []
            for sgn in (1., -1.):
                bb = b.copy()
                bb[i] += sgn * h
                sol.append(s.solve(torch.from_numpy(bb).cuda()).cpu().numpy())
            fdb = _fd_product(sol[0], sol[1], xt, h)
            assert abs(fdb - g_b[i]) <= 1e-4 * np.abs(g_b).max()
    finally:
        s.close()


def test_adjoint_solve_matches_the_oracle_lu():
    """Nonsymmetric case: the transposed device solve (lambda = A^-T g) equals
    the oracle's transposed solve on the same factors' structure."""
    torch = _torch()
    from pastix_amd.autograd import SparseSolver
    grid = (12, 11, 10)
    A = problems.convection_diffusion_27(*grid)
    n = A.shape[0]
    s = SparseSolver(A, symmetric=False, grid=grid, refine=False)
    try:
        C = sp.csc_matrix(A)
        C.sort_indices()
        s.factor(torch.from_numpy(C.data.copy()))
        g = problems.uniform(n, seed=6)
        lam = s.solve(torch.from_numpy(g).cuda(), transpose=True).cpu().numpy()
        P = Problem(A, grid, pa.factotype.LU, analyze_only=True)
        try:
            L, U, _ = P.oracle_factors()
            lp = orc.solve(P.S, 2, L, U, g[P.invp], trans=1)
            ref = np.empty(n)
            ref[P.invp] = lp
        finally:
            P.close()
        assert np.abs(lam - ref).max() <= 1e-12 * np.abs(ref).max()
        assert np.abs(A.T @ lam - g).max() <= 1e-10 * np.abs(g).max()
    finally:
        s.close()


def test_gradient_matches_one_assembled_from_oracle_solves():
    """dL/dvalues and dL/db of the device autograd path against a gradient
    assembled from the CPU oracle's solves on the same factorization
    structure: x = A^-1 b, lambda = A^-T (x - x_t), dL/dA_ij = -lambda_i x_j on
    the stored pattern, dL/db = lambda (nonsymmetric LU case)."""
    torch = _torch()
    from pastix_amd.autograd import SparseSolver, solve
    grid = (11, 10, 9)
    A = sp.csc_matrix(problems.convection_diffusion_27(*grid))
    A.sort_indices()
    n = A.shape[0]
    b = problems.uniform(n, seed=7)
    xt = problems.uniform(n, seed=8)
    s = SparseSolver(A, symmetric=False, grid=grid, refine=False)
    try:
        vals = torch.from_numpy(A.data.copy()).cuda().requires_grad_(True)
        bt = torch.from_numpy(b).cuda().requires_grad_(True)
        x = solve(s, vals, bt)
        (0.5 * ((x - torch.from_numpy(xt).cuda()) ** 2).sum()).backward()
        g_vals = vals.grad.cpu().numpy()
        g_b = bt.grad.cpu().numpy()
    finally:
        s.close()
    P = Problem(A, grid, pa.factotype.LU, analyze_only=True)
    try:
        L, U, _ = P.oracle_factors()
        xo = np.empty(n)
        xo[P.invp] = orc.solve(P.S, 2, L, U, b[P.invp])
        lam = np.empty(n)
        lam[P.invp] = orc.solve(P.S, 2, L, U, (xo - xt)[P.invp], trans=1)
    finally:
        P.close()
    cols = np.repeat(np.arange(n), np.diff(A.indptr))
    ref = -lam[A.indices] * xo[cols]
    assert np.abs(g_b - lam).max() <= 1e-11 * np.abs(lam).max()
    assert np.abs(g_vals - ref).max() <= 1e-11 * np.abs(ref).max()
