"""Chebyshev collocation operators of the glow-discharge x-space: the reference's
collocation_op.cheb_collocation_1d (BESolver/python/collocation_op.py:8-51), which
mesh.mesh(..., CHEBYSHEV_COLLOC) hands to the glow solver as xcoord, D1 and D2.

Host set-up of an Np x Np table, done once.  The points are the Chebyshev-Gauss-Lobatto
points xp_k = -cos(pi k / (Np-1)) in ascending order (the reference's).  The first-derivative
matrix is the closed form of polynomial interpolation on those points,
D_ij = (c_i / c_j) (-1)^(i+j) / (x_i - x_j) for i != j (c = 2 at both ends, 1 inside), with
the diagonal set so each row sums to zero (the derivative of a constant).  The reference
builds the same matrices from the Chebyshev Vandermonde matrix and its inverse; both are
exact on polynomials of degree < Np and agree to rounding.  Lp = Dp Dp is the second
derivative (exact on the same space); LpD has identity boundary rows for the Dirichlet
Poisson problem.
"""
import numpy as np


class cheb_collocation_1d:
    def __init__(self, Np):
        if Np < 3:
            raise ValueError("Np >= 3 Chebyshev points are needed")
        self.Np = Np
        self.deg = Np - 1
        N = self.deg
        self.xp = -np.cos(np.pi * np.arange(Np) / N)
        c = np.ones(Np)
        c[0] = c[-1] = 2.0
        c *= (-1.0) ** np.arange(Np)
        dx = self.xp[:, None] - self.xp[None, :]
        D = np.outer(c, 1.0 / c) / (dx + np.eye(Np))
        np.fill_diagonal(D, 0.0)
        D[np.diag_indices(Np)] = -D.sum(axis=1)
        self.Dp = D
        self.Lp = D @ D
        self.LpD = np.identity(Np)
        self.LpD[1:-1, :] = self.Lp[1:-1, :]
        self.LpD_inv = np.linalg.solve(self.LpD, np.eye(Np))
