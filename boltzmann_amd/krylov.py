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
import gc

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


def _ptr(t, offset=0):
    """Device address of element `offset` of a contiguous tensor of any element type."""
    assert t.is_cuda and t.is_contiguous()
    return ctypes.c_void_p(t.data_ptr() + t.element_size() * offset)


class capture_guard:
    """Collects dead Python reference cycles before a stream capture and keeps the cyclic
    garbage collector off during it.  A dead cycle can hold captured graphs or device memory
    (a solver and its GMRES object referring to each other); if the collector frees them in
    the middle of another capture, the graph's teardown and the freeing of its memory pool
    are device calls that a capturing stream does not allow, and the process aborts.  This
    torch release no longer collects at the start of a capture by default."""

    def __enter__(self):
        gc.collect()
        self._was = gc.isenabled()
        gc.disable()
        return self

    def __exit__(self, *exc):
        if self._was:
            gc.enable()
        return False


class DeviceGMRES:
    """Restarted, right-preconditioned GMRES with the whole Arnoldi step on the device, for an
    operator and a preconditioner given as launchers `matvec(src, dst)` and
    `precond(src, dst)` (device addresses of flat float64 vectors of length n; each writes
    dst only).

    The method of utils.gmres_pcr (utils.py:1172-1268: right preconditioning, restart,
    tol = max(atol, rtol |b|)) with the host out of the inner loop: the stop test of a step
    is the least-squares residual |g_(i+1)| of the Givens-rotated Hessenberg system (the
    true residual in exact arithmetic, bz_gmres_givens), and the true residual is evaluated
    once per cycle, at its end, before the solution is accepted.  Per step: preconditioner,
    operator, classical Gram-Schmidt (bz_multi_dot, bz_vt_axpy; two passes by default, one
    with gs_passes=1), the norm and scaling, one rotation.  After the first cycle the steps are replayed from HIP graphs of `chunk`
    steps; the host reads the device stop flag between graphs.  The launchers' scalar
    arguments and buffers are captured with the graphs, so they must stay the same between
    solves (inputs are copied into fixed buffers)."""

    def __init__(self, n, restart, dev, matvec, precond, chunk=4, use_graph=True, gs_passes=2):
        f64, i32 = torch.float64, torch.int32
        self.n, self.m, self.chunk = int(n), int(restart), int(chunk)
        self.gs_passes = int(gs_passes)                             # classical Gram-Schmidt passes
        self.matvec, self.precond = matvec, precond
        self.use_graph = use_graph
        self.nblk = max(1, min(1024, (self.n + 4095) // 4096))
        m = self.m
        # zero-filled: rows past the last step of a cycle enter the solution update with zero
        # coefficients, which must not meet uninitialised NaN patterns
        self.V = torch.zeros((m + 1, self.n), dtype=f64, device=dev)
        self.H = torch.zeros((m + 2, m + 1), dtype=f64, device=dev)
        self.ldh = m + 1
        self.partial = torch.empty(((m + 1) * self.nblk,), dtype=f64, device=dev)
        self.h = torch.zeros(m + 2, dtype=f64, device=dev)
        self.cs = torch.zeros(m + 1, dtype=f64, device=dev)
        self.sn = torch.zeros(m + 1, dtype=f64, device=dev)
        self.g = torch.zeros(m + 2, dtype=f64, device=dev)
        self.y = torch.zeros(m + 1, dtype=f64, device=dev)
        self.tol = torch.zeros(1, dtype=f64, device=dev)
        self.scal = torch.zeros(4, dtype=f64, device=dev)
        self.flags = torch.zeros(2, dtype=i32, device=dev)          # done, steps of the cycle
        self.w = torch.empty(self.n, dtype=f64, device=dev)
        self.t = torch.empty(self.n, dtype=f64, device=dev)
        self.x = torch.empty(self.n, dtype=f64, device=dev)
        self.b = torch.empty(self.n, dtype=f64, device=dev)
        self.r = torch.empty(self.n, dtype=f64, device=dev)
        self.graphs = None

    def _dot(self, V, nv, u, out, ostride=1, accumulate=0, take_sqrt=0, out2=None):
        _lib.call("bz_multi_dot", _at(V), self.n, nv, _at(u), self.n, self.partial, self.nblk, out, ostride,
                  accumulate, take_sqrt, out2)

    def _step(self, i):
        n, ldh = self.n, self.ldh
        self.precond(_at(self.V, i * n), _at(self.t))
        self.matvec(_at(self.t), _at(self.w))
        for rep in range(self.gs_passes):
            self._dot(self.V, i + 1, self.w, _at(self.H, i), ostride=ldh, accumulate=rep, out2=self.h)
            _lib.call("bz_vt_axpy", _at(self.V), n, i + 1, self.h, 1, -1.0, self.w, self.w, n)
        self._dot(self.w, 1, self.w, _at(self.H, (i + 1) * ldh + i), take_sqrt=1)
        _lib.call("bz_scale_inv", self.w, _at(self.H, (i + 1) * ldh + i), _at(self.V, (i + 1) * n), n)
        _lib.call("bz_gmres_givens", self.H, ldh, i, self.cs, self.sn, self.g, self.tol, _ptr(self.flags, 0),
                  _ptr(self.flags, 1))

    def _capture(self):
        with capture_guard():
            self.graphs = []
            for i0 in range(0, self.m, self.chunk):
                gr = torch.cuda.CUDAGraph()
                with torch.cuda.graph(gr):
                    for i in range(i0, min(self.m, i0 + self.chunk)):
                        self._step(i)
                self.graphs.append(gr)

    def _residual(self, out_scal):
        """r = b - A P x, |r| into scal[out_scal]; leaves P x in t."""
        self.precond(_at(self.x), _at(self.t))
        self.matvec(_at(self.t), _at(self.r))
        torch.sub(self.b, self.r, out=self.r)
        self._dot(self.r, 1, self.r, _at(self.scal, out_scal), take_sqrt=1)

    def solve(self, b, atol, rtol, max_cycles, x0=None, callback=None):
        """Returns (P x, total steps); raises RuntimeError after max_cycles cycles."""
        self.b.copy_(b.reshape(-1))
        if x0 is None:
            self.x.zero_()
        else:
            self.x.copy_(x0.reshape(-1))
        self._dot(self.b, 1, self.b, _at(self.scal, 1), take_sqrt=1)
        self._residual(0)                                  # r = b - A P x0, |r| -> scal[0]
        iters = 0
        for cyc in range(max_cycles):
            # r and |r| are those of the current x: computed above, or at the end of the
            # previous cycle (where the host found |r| >= tol), so only the first cycle reads
            # the start test back
            self.H.zero_()
            _lib.call("bz_gmres_cycle_start", self.scal, float(atol), float(rtol), self.m, self.g, self.tol,
                      _ptr(self.flags, 0), _ptr(self.flags, 1))
            if cyc == 0 and int(self.flags[0].item()):
                return self.t.clone(), iters
            _lib.call("bz_scale_inv", self.r, self.scal, _at(self.V, 0), self.n)
            if self.graphs is None or not self.use_graph:
                # eager cycle (the first one loads and runs every kernel once), then capture
                for i in range(self.m):
                    self._step(i)
                    if (i + 1) % self.chunk == 0 and int(self.flags[0].item()):
                        break
                if self.use_graph:
                    self._capture()
            else:
                for gr in self.graphs:
                    gr.replay()
                    if int(self.flags[0].item()):
                        break
            _lib.call("bz_gmres_back_solve", self.H, self.ldh, self.g, _ptr(self.flags, 1), self.y, self.m)
            _lib.call("bz_vt_axpy", _at(self.V), self.n, self.m, self.y, 1, 1.0, self.x, self.x, self.n)
            self._residual(2)
            # one read-back per cycle: the true residual, |b|, the tolerance and the step count
            rn, bn, tol, k = torch.stack([self.scal[2], self.scal[1], self.tol[0],
                                          self.flags[1].to(torch.float64)]).tolist()
            iters += int(k)
            if callback is not None:
                callback(rn / bn if bn > 0 else 0.0)
            if rn < tol:
                return self.t.clone(), iters
            self.scal[0].copy_(self.scal[2])              # the next cycle starts from this r
        raise RuntimeError(f"GMRES did not converge in {max_cycles} cycles of {self.m}")
