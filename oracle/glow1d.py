"""Oracle: 1D3V ordinates, ordinate <-> spherical-harmonic transforms and the x-advection
half step (FP64, numpy).  TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates BESolver/python/glowdischarge_boltzmann_1d.py:
  * ordinates: half-range Gauss-Legendre in cos(theta) on [-1,0] and [0,1], theta =
    arccos (:180-185);
  * po2sh[l, j] = Y_l0(theta_j) w_j 2 pi, psh2o = Y_l0(theta_j)^T (:330-334) and
    ords_to_sph / sph_to_ords (:1032-1048);
  * the x-advection operator adv_x = M^-1 * scale * int v^3 B_p B_k (:369-472, scale =
    vth tau / L), its eigen decomposition, the backward-Euler upwind shift operators
    (:718-790, first-order upwind differences) and step_bte_x (:1220-1287).

Pinned by tests/test_oracle_glow1d.py: po2sh psh2o = identity on the span of the
ordinates up to lmax (exact quadrature), and step_bte_x against a dense solve of
(I + dt cos(theta_j) adv_x (x) D) per ordinate with the inflow rows zeroed.
"""
import math

import numpy as np

from . import operators as op


def ordinates(Nvt):
    """(theta_j, w_j): Nvt/2 Gauss points on each half of cos(theta) in [-1, 1]."""
    assert Nvt % 2 == 0
    gx, gw = np.polynomial.legendre.leggauss(Nvt // 2)
    x = np.append(0.5 * gx - 0.5, 0.5 * gx + 0.5)
    w = np.append(0.5 * gw, 0.5 * gw)
    return np.arccos(x), w


def sph_values(lmax, theta):
    """Y[l, j] = Y_l0(theta_j) (real, orthonormal on the sphere)."""
    mu = np.cos(theta)
    Y = np.zeros((lmax + 1, len(theta)))
    for l in range(lmax + 1):
        c = np.zeros(l + 1)
        c[l] = 1.0
        Y[l] = math.sqrt((2 * l + 1) / (4 * math.pi)) * np.polynomial.legendre.legval(mu, c)
    return Y


def sph_ops(lmax, theta, w):
    Y = sph_values(lmax, theta)
    return Y * w[None, :] * 2 * math.pi, Y.T.copy()


def ords_to_sph(po2sh, v, num_p, Nvt, Nx):
    return np.einsum("li,rix->rlx", po2sh, v.reshape((num_p, Nvt, Nx))).reshape((-1, Nx))


def sph_to_ords(psh2o, v, num_p, num_sh, Nx):
    return np.einsum("li,rix->rlx", psh2o, v.reshape((num_p, num_sh, Nx))).reshape((-1, Nx))


def upwind_first_order(x):
    """Two-point one-sided first derivatives: D_L (backward, row 0 zero) and D_R
    (forward, last row zero)."""
    N = len(x)
    DL = np.zeros((N, N))
    DR = np.zeros((N, N))
    for i in range(1, N):
        h = x[i] - x[i - 1]
        DL[i, i], DL[i, i - 1] = 1.0 / h, -1.0 / h
    for i in range(N - 1):
        h = x[i + 1] - x[i]
        DR[i, i], DR[i, i + 1] = -1.0 / h, 1.0 / h
    return DL, DR


def adv_x_operator(g: op.Grid, scale):
    """M^-1 * scale * [int v^3 B_p B_k] on the radial basis (exact Gauss rule)."""
    x, w = g.quad(g.sp_order + 8)
    V = g.V(x)
    S3 = (V * (w * x ** 3)) @ V.T
    M = (V * (w * x ** 2)) @ V.T
    return np.linalg.solve(M, scale * S3)


class XStep:
    """initialize_bte_adv_x (upwind branch) + step_bte_x."""

    def __init__(self, adv_x, xgrid, theta, dt):
        self.d, self.Q = np.linalg.eig(adv_x)
        assert np.abs(self.d.imag).max() == 0 and np.abs(self.Q.imag).max() == 0
        self.d, self.Q = self.d.real, self.Q.real
        self.Qinv = np.linalg.inv(self.Q)
        self.theta = theta
        Nr, Nvt, Nx = len(self.d), len(theta), len(xgrid)
        self.shape = (Nr, Nvt, Nx)
        DL, DR = upwind_first_order(xgrid)
        self.lower = theta <= 0.5 * np.pi
        self.diag = np.zeros((Nr, Nvt, Nx))
        self.sdiag = np.zeros((Nr, Nvt, Nx - 1))
        for j in range(Nvt):
            D, off = (DL, -1) if self.lower[j] else (DR, 1)
            for i in range(Nr):
                a = dt * self.d[i] * np.cos(theta[j])
                self.diag[i, j] = 1.0 + a * np.diagonal(D, 0)
                self.sdiag[i, j] = a * np.diagonal(D, off)

    def step(self, v):
        Nr, Nvt, Nx = self.shape
        z = (self.Qinv @ v.reshape(Nr, Nvt * Nx)).reshape(Nr, Nvt, Nx)
        z[:, self.lower, 0] = 0.0
        z[:, ~self.lower, -1] = 0.0
        # every (mode, ordinate) bidiagonal system along x at once: forward substitution for
        # the ordinates with inflow at x_0, backward for those with inflow at x_end
        y = np.zeros_like(z)
        lo, up = self.lower, ~self.lower
        dg, sd = self.diag, self.sdiag
        y[:, lo, 0] = z[:, lo, 0] / dg[:, lo, 0]
        for k in range(1, Nx):
            y[:, lo, k] = (z[:, lo, k] - sd[:, lo, k - 1] * y[:, lo, k - 1]) / dg[:, lo, k]
        y[:, up, -1] = z[:, up, -1] / dg[:, up, -1]
        for k in range(Nx - 2, -1, -1):
            y[:, up, k] = (z[:, up, k] - sd[:, up, k] * y[:, up, k + 1]) / dg[:, up, k]
        return (self.Q @ y.reshape(Nr, Nvt * Nx)).reshape(Nr * Nvt, Nx)
