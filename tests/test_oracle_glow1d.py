"""Pins of oracle/glow1d.py (1D3V ordinates, ordinate <-> harmonic transforms, x-step).

* po2sh psh2o acts as the identity on the harmonics up to lmax: the half-range
  Gauss rule integrates Y_l Y_l' exactly (degree 2 lmax < Nvt);
* step_bte_x (eigen coordinates, per-mode bidiagonal sweeps) equals a dense solve of
  (I + dt cos(theta_j) adv_x (x) D_dir) per ordinate with the inflow rows of the
  right-hand side zeroed - no eigen decomposition on that side;
* adv_x is the Galerkin form M^-1 scale int v^3 B_p B_k: the exact sympy integrals.
"""
import math
import os

import numpy as np
import pytest

from oracle import glow1d as gl
from oracle import operators as op

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "bspline_exact.npz"))


@pytest.mark.parametrize("lmax,Nvt", [(1, 8), (2, 16), (4, 32)])
def test_ordinate_transforms_invert_on_harmonics(lmax, Nvt):
    th, w = gl.ordinates(Nvt)
    assert np.all(np.diff(np.cos(th)) > 0) and abs(w.sum() - 2.0) < 1e-14
    po2sh, psh2o = gl.sph_ops(lmax, th, w)
    np.testing.assert_allclose(po2sh @ psh2o, np.eye(lmax + 1), atol=1e-13)
    f = 1 + np.cos(th)
    np.testing.assert_allclose(psh2o @ (po2sh @ f), f, atol=1e-13)


def test_adv_x_operator_exact():
    t = GOLD["uniform_knots"]
    g = op.Grid(t, len(t) - 4, 0, 5, 1.0)
    A = gl.adv_x_operator(g, 2.5)
    np.testing.assert_allclose(GOLD["uniform_M"] @ A, 2.5 * GOLD["uniform_S3"], atol=1e-12)


def test_xstep_equals_dense_solve():
    t = GOLD["uniform_knots"]
    g = op.Grid(t, len(t) - 4, 0, 5, 1.0)
    adv_x = gl.adv_x_operator(g, 0.7)
    Nvt, Nx, dt = 6, 11, 0.13
    th, _ = gl.ordinates(Nvt)
    xg = -np.cos(np.pi * np.arange(Nx) / (Nx - 1))
    xs = gl.XStep(adv_x, xg, th, dt)
    rng = np.random.default_rng(0)
    Nr = adv_x.shape[0]
    v = rng.normal(size=(Nr * Nvt, Nx))
    out = xs.step(v).reshape(Nr, Nvt, Nx)
    DL, DR = gl.upwind_first_order(xg)
    V = v.reshape(Nr, Nvt, Nx)
    for j in range(Nvt):
        D = DL if th[j] <= 0.5 * np.pi else DR
        A = np.eye(Nr * Nx) + dt * np.cos(th[j]) * np.kron(adv_x, D)
        rhs = V[:, j, :].copy()
        rhs[:, 0 if th[j] <= 0.5 * np.pi else -1] = 0.0
        ref = np.linalg.solve(A, rhs.reshape(-1)).reshape(Nr, Nx)
        np.testing.assert_allclose(out[:, j, :], ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
