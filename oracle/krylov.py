"""Oracle: restarted right-preconditioned GMRES (utils.gmres_pcr) and the 1D3V
backward-Euler v-space step (FP64, numpy).  TEST INFRASTRUCTURE ONLY (see
oracle/__init__.py).

Restates
  * utils.gmres_pcr (BESolver/python/utils.py:1172-1268): modified Gram-Schmidt
    Arnoldi, least squares of the Hessenberg matrix after every step, true residual
    check after every step;
  * step_bte_v, ts_type "BE" (glowdischarge_boltzmann_1d.py:1476-1592): the operator
    I - dt tau (n0 np0 (C + Tg C_T) + E(x) A) is block diagonal in x, so its exact
    solution is one dense solve per point, which is what the tests compare with.

Pinned by tests/test_oracle_krylov.py against numpy.linalg.solve.
"""
import numpy as np


def gmres_pcr(A, b, x0=None, atol=1e-20, rtol=1e-10, maxiter=None, restart=None, M=None):
    """A, M: callables (matvec) on flat vectors."""
    n = b.shape[0]
    b_norm = np.linalg.norm(b)
    if b_norm == 0:
        return b, 0
    if maxiter is None:
        maxiter = n * 10
    if restart is None:
        restart = maxiter
    psolve = M if M is not None else (lambda v: v)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    tol = max(atol, rtol * b_norm)
    it = 0
    xn = x
    while it < maxiter:
        V = np.empty((n, restart + 1))
        H = np.zeros((restart + 2, restart + 1))
        e = np.zeros(restart + 2)
        mx = psolve(x)
        r = b - A(mx)
        r_norm = np.linalg.norm(r)
        if r_norm < tol:
            return mx, 0
        V[:, 0] = r / r_norm
        e[0] = r_norm
        for i in range(restart):
            u = A(psolve(V[:, i]))
            for j in range(i + 1):
                H[j, i] = np.dot(V[:, j], u)
                u = u - H[j, i] * V[:, j]
            H[i + 1, i] = np.linalg.norm(u)
            if H[i + 1, i] != 0:
                V[:, i + 1] = u / H[i + 1, i]
            y = np.linalg.lstsq(H[0:i + 2, 0:i + 1], e[0:i + 2], rcond=None)[0]
            xn = x + V[:, 0:i + 1] @ y
            rn = b - A(psolve(xn))
            if np.linalg.norm(rn) < tol:
                return psolve(xn), 0
            it += 1
        x = xn
    return psolve(xn), it + 1


def vstep_exact(K0, A, E, dt, tau, u):
    """Per-point dense solve of (I - dt tau (K0 + E_x A)) y_x = u_x; u (dof, Nx)."""
    dof = K0.shape[0]
    I = np.eye(dof)
    L = I[None] - dt * tau * (K0[None] + np.asarray(E)[:, None, None] * A[None])
    return np.linalg.solve(L, u.T[:, :, None])[:, :, 0].T
