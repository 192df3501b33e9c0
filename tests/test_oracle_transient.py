"""Pins of the oracle's transient (backward-Euler) 0D path (bte_0d3v_batched.py:59-126,
981-1161 restated in oracle/solver0d.py):
  * the new state satisfies the backward-Euler equation Q^T (u_new - u) = dt r(u_new)
    in the Q basis, checked by evaluating the residual directly (not through the
    Newton loop), and stays on u.x = 1;
  * a steady state is a fixed point of the step, for several step sizes.
"""
import numpy as np
import pytest

from helpers import instance_params, oracle_batched
from oracle import solver0d as s0


@pytest.fixture(scope="module")
def case():
    ops, g, procs = oracle_batched(24, 1)
    prm = instance_params(3, seed=11)
    return ops, prm


@pytest.mark.parametrize("c", [1e-11, 1e-10])
def test_step_solves_backward_euler_equation(case, c):
    ops, prm = case
    u = ops.initial(3)
    u = u / (ops.u @ u)
    un, info = s0.transient_step(ops, prm, u, c, 0.0, 1e-13, 60)
    assert info["status"]
    res, _ = s0.batched_res_jac(ops, prm)
    G = ops.Q.T @ (un - u) - c * res(un)
    assert np.linalg.norm(G, axis=0).max() <= 1e-12 * np.linalg.norm(ops.Q.T @ (un - u), axis=0).max()
    np.testing.assert_allclose(ops.u @ un, 1.0, rtol=1e-12)


@pytest.mark.parametrize("c", [1e-10, 5e-9])
def test_steady_state_is_fixed_point(case, c):
    ops, prm = case
    x, q, info = s0.steady_state_solve_each(ops, prm, ops.initial(3), 0.0, 1e-14, 40, jacobian="exact")
    x = x / (ops.u @ x)
    un, info = s0.transient_step(ops, prm, x, c, 0.0, 1e-13, 50)
    assert np.abs(un - x).max() <= 1e-9 * np.abs(x).max()
