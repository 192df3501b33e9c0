"""Pins of oracle/krylov.py: GMRES (utils.gmres_pcr restated) converges to the
numpy.linalg.solve solution, with and without restarts and with a right
preconditioner; a lucky breakdown (exact Krylov space) returns the exact solution."""
import numpy as np
import pytest

from oracle import krylov as kr


@pytest.mark.parametrize("restart,precond", [(None, False), (5, False), (8, True)])
def test_gmres_matches_dense_solve(restart, precond):
    rng = np.random.default_rng(3)
    n = 60
    A = np.eye(n) * 4 + rng.normal(size=(n, n)) * 0.3
    b = rng.normal(size=n)
    M = None
    if precond:
        Minv = np.linalg.inv(np.diag(np.diag(A)))
        M = lambda v: Minv @ v
    x, info = kr.gmres_pcr(lambda v: A @ v, b, rtol=1e-12, restart=restart, maxiter=500, M=M)
    assert info == 0
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-10)


def test_gmres_exact_krylov_space():
    A = np.diag([1.0, 2.0, 3.0, 4.0])
    b = np.array([1.0, 1.0, 0.0, 0.0])
    x, info = kr.gmres_pcr(lambda v: A @ v, b, rtol=1e-14)
    assert info == 0
    np.testing.assert_allclose(x, [1.0, 0.5, 0.0, 0.0], atol=1e-14)
