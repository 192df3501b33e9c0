"""Restarted, right-preconditioned GMRES on the device.

Mirrors utils.gmres_pcr (BESolver/python/utils.py:1172-1268): same arguments, same
right preconditioning (A M^-1 y = b, x = M^-1 y), the least-squares problem of the
Hessenberg matrix solved after every Arnoldi step, a true-residual check b - A M^-1 x_n
after every step, tol = max(atol, rtol |b|), maxiter tested between restart cycles only (a
cycle always runs to its end), return value (solution, info) with info 0 on convergence and
the iteration count otherwise.

Design: the Krylov basis lives in HBM as rows of one (restart+1, n) array; every
Arnoldi step is two block dot products and two block updates (classical Gram-Schmidt
applied twice - the reference uses modified Gram-Schmidt; both give an orthonormal
basis to FP64 rounding), one norm and one scaling, all in HIP kernels
(bz_krylov.hip).  The Hessenberg least-squares problem is solved on the device by
Givens rotations.  The host reads one scalar per step (the true residual norm), as
the reference does.
"""
import ctypes

import torch

from . import _lib


def _at(t, offset=0):
    """Device address of element `offset` of a contiguous tensor (for strided views the
    kernels take an explicit stride)."""
    assert t.is_cuda and t.is_contiguous()
    return ctypes.c_void_p(t.data_ptr() + 8 * offset)


class _Work:
    def __init__(self, n, restart, dev):
        self.n = n
        self.nblk = max(1, min(1024, (n + 4095) // 4096))
        self.V = torch.empty((restart + 1, n), dtype=torch.float64, device=dev)
        self.H = torch.zeros((restart + 2, restart + 1), dtype=torch.float64, device=dev)
        self.ldh = restart + 1
        self.partial = torch.empty(((restart + 1) * self.nblk,), dtype=torch.float64, device=dev)
        self.h = torch.zeros(restart + 2, dtype=torch.float64, device=dev)
        self.scal = torch.zeros(4, dtype=torch.float64, device=dev)
        self.lsq_R = torch.empty(((restart + 2) * (restart + 1),), dtype=torch.float64, device=dev)
        self.lsq_y = torch.empty((2 * restart + 2,), dtype=torch.float64, device=dev)
        self.u = torch.empty(n, dtype=torch.float64, device=dev)

    def dot(self, V, nv, u, out, ostride=1, accumulate=0, take_sqrt=0, out2=None):
        _lib.call("bz_multi_dot", _at(V), self.n, nv, _at(u), self.n, self.partial, self.nblk, out, ostride, accumulate,
                  take_sqrt, out2)

    def norm(self, x):
        """|x| into scal[0] (device); returns that 1-element view."""
        self.dot(x, 1, x, _at(self.scal), take_sqrt=1)
        return self.scal[0:1]


def gmres_pcr(A, b, x0=None, atol=1e-20, rtol=1e-10, maxiter=None, restart=None, M=None, callback=None):
    """A, M: objects with .matvec(x) on flat device vectors (or callables); b a flat
    float64 device vector.  Returns (x, info)."""
    n = b.numel()
    if n == 0:
        return torch.empty_like(b), 0
    matvec = A.matvec if hasattr(A, "matvec") else A
    psolve = (M.matvec if hasattr(M, "matvec") else M) if M is not None else (lambda v: v)
    if maxiter is None:
        maxiter = n * 10
    if restart is None:
        restart = maxiter
    restart = int(restart)
    w = _Work(n, restart, b.device)
    b = b.contiguous()
    b_norm = float(w.norm(b).item())
    if b_norm == 0:
        return b, 0
    tol = max(atol, rtol * b_norm)
    x = torch.zeros_like(b) if x0 is None else x0.reshape(-1).contiguous().clone()
    V, H, ldh = w.V, w.H, w.ldh
    beta = torch.zeros(1, dtype=torch.float64, device=b.device)
    it = 0
    xn = x
    while it < maxiter:
        H.zero_()
        mx = psolve(x)
        r = (b - matvec(mx)).contiguous()
        beta.copy_(w.norm(r))
        if float(beta.item()) < tol:
            return mx, 0
        _lib.call("bz_scale_inv", r, beta, _at(V, 0), n)
        for i in range(restart):
            w.u.copy_(matvec(psolve(V[i])))
            # H[0:i+1, i] = V^T u, applied twice (classical Gram-Schmidt, re-orthogonalised)
            for rep in range(2):
                w.dot(V, i + 1, w.u, _at(H, i), ostride=ldh, accumulate=rep, out2=w.h)
                _lib.call("bz_vt_axpy", _at(V), n, i + 1, w.h, 1, -1.0, w.u, w.u, n)
            w.dot(w.u, 1, w.u, _at(H, (i + 1) * ldh + i), take_sqrt=1)
            _lib.call("bz_scale_inv", w.u, _at(H, (i + 1) * ldh + i), _at(V, (i + 1) * n), n)
            _lib.call("bz_hess_lstsq", H, ldh, i + 1, beta, w.lsq_R, w.lsq_y)
            xn = torch.empty_like(x)
            _lib.call("bz_vt_axpy", _at(V), n, i + 1, w.lsq_y, 1, 1.0, x, xn, n)
            rn = (b - matvec(psolve(xn))).contiguous()
            rn_norm = float(w.norm(rn).item())
            if callback is not None:
                callback(rn_norm / b_norm)
            if rn_norm < tol:
                return psolve(xn), 0
            it += 1
        x = xn
    return psolve(xn), it + 1
