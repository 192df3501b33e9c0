"""1D3V velocity-space pieces of the glow-discharge Boltzmann step, on the MI355X.

Mirrors the v-space side of glow1d_boltzmann (BESolver/python/glowdischarge_boltzmann_1d.py):
  * ordinates theta_j and weights (:180-185), po2sh / psh2o (:330-334),
    ords_to_sph / sph_to_ords (:1032-1048)        -> bz_slab_left_mult
  * the x-advection half step: adv_x = M^-1 (vth tau / L) [int v^3 B_p B_k] (:369-472),
    its eigen decomposition, the upwind backward-Euler shift operators
    (:718-790, USE_BE_UPW_FD) and step_bte_x (:1220-1287)
                                                   -> rocBLAS GEMMs with Q^-1, Q and one
                                                      bz_xstep_bidiag_solve launch
The state is v (Nr*Nvt, Nx): radial B-spline coefficient r, ordinate j, point x, with x
fastest (the reference's layout).  Host work is set-up only (the eigen decomposition
of one Nr x Nr matrix, as in the reference); every per-step operation runs on the GPU.
"""
import math

import numpy as np
import torch

from . import _lib


def ordinates(Nvt):
    """glowdischarge_boltzmann_1d.py:180-185: Nvt/2 Gauss-Legendre points on each half
    of cos(theta) in [-1, 1]; returns (theta, weights)."""
    if Nvt % 2:
        raise ValueError("Nvt must be even")
    gx, gw = np.polynomial.legendre.leggauss(Nvt // 2)
    mu = np.append(0.5 * gx - 0.5, 0.5 * gx + 0.5)
    return np.arccos(mu), np.append(0.5 * gw, 0.5 * gw)


def sph_ordinate_ops(lmax, theta, w):
    """(po2sh (ns, Nvt), psh2o (Nvt, ns)) with Y_l0(theta) real orthonormal (:330-334)."""
    mu = np.cos(theta)
    Y = np.zeros((lmax + 1, len(theta)))
    for l in range(lmax + 1):
        c = np.zeros(l + 1)
        c[l] = 1.0
        Y[l] = math.sqrt((2 * l + 1) / (4 * math.pi)) * np.polynomial.legendre.legval(mu, c)
    return Y * w[None, :] * 2 * math.pi, np.ascontiguousarray(Y.T)


def slab_left_mult(P, x, nr, nx):
    """out[r] = P @ x[r] for the (k, nx) slabs of x (nr*k, nx)."""
    m, k = P.shape
    assert x.shape == (nr * k, nx) and x.dtype == torch.float64
    out = torch.empty((nr * m, nx), dtype=torch.float64, device=x.device)
    _lib.call("bz_slab_left_mult", P.contiguous(), m, k, x.contiguous(), nr, nx, out)
    return out


class Ordinates:
    """Ordinate <-> spherical-harmonic transforms on the device."""

    def __init__(self, num_p, lmax, Nvt, device):
        self.num_p, self.num_sh, self.Nvt = num_p, lmax + 1, Nvt
        self.theta, self.w = ordinates(Nvt)
        po2sh, psh2o = sph_ordinate_ops(lmax, self.theta, self.w)
        self.po2sh = torch.from_numpy(po2sh).to(device)
        self.psh2o = torch.from_numpy(psh2o).to(device)

    def ords_to_sph(self, x):
        """(num_p*Nvt, Nx) -> (num_p*num_sh, Nx) (:1032-1039)."""
        return slab_left_mult(self.po2sh, x, self.num_p, x.shape[1])

    def sph_to_ords(self, x):
        """(num_p*num_sh, Nx) -> (num_p*Nvt, Nx) (:1041-1048)."""
        return slab_left_mult(self.psh2o, x, self.num_p, x.shape[1])


def upwind_first_order(x):
    """Two-point one-sided first derivatives on the points x: D_L (backward differences,
    first row zero) and D_R (forward differences, last row zero)."""
    N = len(x)
    h = np.diff(x)
    dl_d = np.zeros(N)
    dl_s = np.zeros(N - 1)
    dr_d = np.zeros(N)
    dr_s = np.zeros(N - 1)
    dl_d[1:] = 1.0 / h
    dl_s[:] = -1.0 / h            # D_L[i, i-1], i = 1..N-1
    dr_d[:-1] = -1.0 / h
    dr_s[:] = 1.0 / h             # D_R[i, i+1], i = 0..N-2
    return (dl_d, dl_s), (dr_d, dr_s)


class XAdvection:
    """initialize_bte_adv_x (upwind finite differences) + step_bte_x on the device.

    vgrid: VGrid of the radial basis; xgrid: the Nx spatial points; Nvt ordinates;
    scale = vth * tau / L; dt: the half-step size the shift operators are built for."""

    def __init__(self, vgrid, xgrid, Nvt, scale, dt):
        dev = vgrid.dev
        self.dev = dev
        self.dt = dt
        Nr = vgrid.npr
        self.Nr, self.Nvt, self.Nx = Nr, Nvt, len(xgrid)
        self.theta, _ = ordinates(Nvt)
        # adv_x = M^-1 scale S3 (radial), exact Gauss rule (the same rule for M and S3)
        S3 = vgrid.radial_bilinear(vgrid.sp_order + 8, 3, 0, 0).cpu().numpy()
        M = vgrid.radial_bilinear(vgrid.sp_order + 8, 2, 0, 0).cpu().numpy()
        adv_x = np.linalg.solve(M, scale * S3)
        d, Q = np.linalg.eig(adv_x)
        if np.abs(np.imag(d)).max() > 0 or np.abs(np.imag(Q)).max() > 0:
            raise ValueError("x-advection operator has complex eigenvalues")
        d, Q = np.real(d), np.real(Q)
        self.adv_x = adv_x
        self.d = d
        self.Q = torch.from_numpy(np.ascontiguousarray(Q)).to(dev)
        self.Qinv = torch.from_numpy(np.ascontiguousarray(np.linalg.inv(Q))).to(dev)
        (dl_d, dl_s), (dr_d, dr_s) = upwind_first_order(np.asarray(xgrid, dtype=np.float64))
        lower = self.theta <= 0.5 * np.pi
        a = dt * d[:, None] * np.cos(self.theta)[None, :]             # (Nr, Nvt)
        diag = np.where(lower[None, :, None], 1.0 + a[:, :, None] * dl_d[None, None, :],
                        1.0 + a[:, :, None] * dr_d[None, None, :])
        sdiag = np.where(lower[None, :, None], a[:, :, None] * dl_s[None, None, :],
                         a[:, :, None] * dr_s[None, None, :])
        self.diag = torch.from_numpy(np.ascontiguousarray(diag.reshape(Nr * Nvt, -1))).to(dev)
        self.sdiag = torch.from_numpy(np.ascontiguousarray(sdiag.reshape(Nr * Nvt, -1))).to(dev)
        self.lower_of = torch.from_numpy(lower.astype(np.int32)).to(dev)

    def step(self, v, dt=None):
        """step_bte_x (:1220-1287): v (Nr*Nvt, Nx) -> advected v."""
        return self._step(v, dt)

    def _step(self, v, dt):
        if dt is not None and dt != self.dt:
            raise AssertionError("x-advection operators were set up for another dt")
        Nr, Nvt, Nx = self.Nr, self.Nvt, self.Nx
        z = torch.empty((Nr * Nvt, Nx), dtype=torch.float64, device=v.device)
        _lib.call("bz_dgemm_nn", self.Qinv, Nr, Nr, v.contiguous(), Nvt * Nx, z)
        y = torch.empty_like(z)
        _lib.call("bz_xstep_bidiag_solve", self.diag, self.sdiag, self.lower_of, z, y, Nr, Nvt, Nx)
        out = torch.empty((Nr * Nvt, Nx), dtype=torch.float64, device=y.device)
        _lib.call("bz_dgemm_nn", self.Q, Nr, Nr, y.contiguous(), Nvt * Nx, out)
        return out


class _DeviceGMRES:
    """The v-step system (I - dt tau (K0 + E_x A)) y = u, right-preconditioned by the E-bin
    block inverses, by restarted GMRES with the whole Arnoldi step on the device.

    The same Krylov method as utils.gmres_pcr (utils.py:1172-1268: right preconditioning,
    x0 = u, restart, tol = max(atol, rtol |b|)), with two changes that remove the host from
    the inner loop: the residual of a step is the least-squares residual |g_(i+1)| of the
    Givens-rotated Hessenberg system (equal to the true residual in exact arithmetic), and
    the true residual is evaluated once per cycle, at its end.  The cycle is replayed from
    HIP graphs of `chunk` steps each (captured on first use); the host reads the device stop
    flag between chunks.  Per step: the binned preconditioner (bz_binned_apply), the operator
    product (bz_vstep_matvec), classical Gram-Schmidt (bz_multi_dot, bz_vt_axpy; one pass,
    gs_passes=2 for two), the norm, the scaling and one Givens rotation (bz_gmres_givens):
    7 launches (9 with two passes), none of them reading back to the host."""

    def __init__(self, vs, nx, chunk=4, gs_passes=1):
        dev = vs.dev
        # classical Gram-Schmidt passes per step: one (the cost of the reference's single
        # modified pass); the v-steps take the same number of steps as with two
        self.vs, self.nx, self.dof = vs, nx, vs.dof
        self.n = n = vs.dof * nx
        self.m = m = vs.rsrt
        self.chunk = chunk
        self.gs_passes = int(gs_passes)
        f64, i32 = torch.float64, torch.int32
        self.nblk = max(1, min(1024, (n + 4095) // 4096))
        self.V = torch.empty((m + 1, n), dtype=f64, device=dev)
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
        self.flags = torch.zeros(2, dtype=i32, device=dev)          # done, kdone
        self.w = torch.empty(n, dtype=f64, device=dev)
        self.t = torch.empty(n, dtype=f64, device=dev)
        self.x = torch.empty(n, dtype=f64, device=dev)
        self.b = torch.empty(n, dtype=f64, device=dev)
        self.r = torch.empty(n, dtype=f64, device=dev)
        self.E = torch.empty(nx, dtype=f64, device=dev)
        self.nbins = vs.PmatE.shape[0]
        self.bins = torch.empty(nx, dtype=i32, device=dev)
        self.gstart = torch.empty(self.nbins, dtype=i32, device=dev)
        self.gcount = torch.empty(self.nbins, dtype=i32, device=dev)
        self.order = torch.empty(nx, dtype=i32, device=dev)
        self.graphs = None
        self.c = vs.dt * vs.tau

    def _done(self):
        return _ptr(self.flags, 0)

    def _precond(self, src, dst, done=None):
        _lib.call("bz_binned_apply", self.vs.PmatE, self.dof, self.nbins, self.gstart, self.gcount, self.order,
                  src, self.nx, dst, done)

    def _matvec(self, src, dst, done=None):
        _lib.call("bz_vstep_matvec", self.vs.K0, self.vs.A, src, self.E, self.c, self.dof, self.nx, dst, done)

    def _dot(self, V, nv, u, out, ostride=1, accumulate=0, take_sqrt=0, out2=None):
        _lib.call("bz_multi_dot", _ptr(V), self.n, nv, _ptr(u), self.n, self.partial, self.nblk, out, ostride,
                  accumulate, take_sqrt, out2)

    def _step(self, i):
        n, ldh = self.n, self.ldh
        done = self._done()
        self._precond(_ptr(self.V, i * n), self.t, done)
        self._matvec(self.t, self.w, done)
        for rep in range(self.gs_passes):
            self._dot(self.V, i + 1, self.w, _ptr(self.H, i), ostride=ldh, accumulate=rep, out2=self.h)
            _lib.call("bz_vt_axpy", _ptr(self.V), n, i + 1, self.h, 1, -1.0, self.w, self.w, n)
        self._dot(self.w, 1, self.w, _ptr(self.H, (i + 1) * ldh + i), take_sqrt=1)
        _lib.call("bz_scale_inv", self.w, _ptr(self.H, (i + 1) * ldh + i), _ptr(self.V, (i + 1) * n), n)
        _lib.call("bz_gmres_givens", self.H, ldh, i, self.cs, self.sn, self.g, self.tol, _ptr(self.flags, 0),
                  _ptr(self.flags, 1))

    def _capture(self):
        from .krylov import capture_guard
        with capture_guard():
            self.graphs = []
            for i0 in range(0, self.m, self.chunk):
                gr = torch.cuda.CUDAGraph()
                with torch.cuda.graph(gr):
                    for i in range(i0, min(self.m, i0 + self.chunk)):
                        self._step(i)
                self.graphs.append(gr)

    def _cycle_start(self, atol, rtol):
        """r = b - L P x, |r| -> scal[0], V[0] = r / |r|, H = 0, g, tol and the flags set."""
        self._precond(self.x, self.t)
        self._matvec(self.t, self.r)
        torch.sub(self.b, self.r, out=self.r)
        self._dot(self.r, 1, self.r, _ptr(self.scal, 0), take_sqrt=1)
        self.H.zero_()
        _lib.call("bz_gmres_cycle_start", self.scal, float(atol), float(rtol), self.m, self.g, self.tol,
                  _ptr(self.flags, 0), _ptr(self.flags, 1))
        _lib.call("bz_scale_inv", self.r, self.scal, _ptr(self.V, 0), self.n)

    def solve(self, u, E, max_cycles=50):
        vs = self.vs
        self.E.copy_(E)
        _lib.call("bz_nearest_bin", self.E, self.nx, vs.Evals, vs.Evals.shape[0], self.bins)
        _lib.call("bz_group_bins", self.bins, self.nx, self.nbins, self.gstart, self.gcount, self.order)
        self.b.copy_(u.reshape(-1))
        self.x.copy_(self.b)                                       # x0 = u, as the reference
        self._dot(self.b, 1, self.b, _ptr(self.scal, 1), take_sqrt=1)
        first = self.graphs is None
        iters = 0
        for cyc in range(max_cycles):
            self._cycle_start(vs.atol, vs.rtol)
            if first:
                # the first cycle runs eagerly (every kernel loaded and run once), then the
                # steps are captured for the later cycles and v-steps
                for i in range(self.m):
                    self._step(i)
                    if (i + 1) % self.chunk == 0 and int(self.flags[0].item()):
                        break
                self._capture()
                first = False
            else:
                for gr in self.graphs:
                    gr.replay()
                    if int(self.flags[0].item()):
                        break
            _lib.call("bz_gmres_back_solve", self.H, self.ldh, self.g, _ptr(self.flags, 1), self.y, self.m)
            _lib.call("bz_vt_axpy", _ptr(self.V), self.n, self.m, self.y, 1, 1.0, self.x, self.x, self.n)
            # the true residual of the cycle's end; one read-back for it, the tolerance and the
            # cycle's step count
            self._precond(self.x, self.t)
            self._matvec(self.t, self.r)
            torch.sub(self.b, self.r, out=self.r)
            self._dot(self.r, 1, self.r, _ptr(self.scal, 2), take_sqrt=1)
            rn, tol, k = torch.stack([self.scal[2], self.tol[0], self.flags[1].to(torch.float64)]).tolist()
            iters += int(k)
            if rn < tol:
                return self.t.reshape(self.dof, self.nx).clone(), iters
        raise RuntimeError(f"v-step GMRES did not converge in {max_cycles} cycles of {self.m}")


def _ptr(t, offset=0):
    import ctypes
    return ctypes.c_void_p(t.data_ptr() + t.element_size() * offset)


def default_pc_fields(num_pc_evals=30):
    """Preconditioner field values of the reference set-up (:2087-2092): +-logspace(1e2, 6e5)."""
    ep = np.logspace(np.log10(1e2), np.log10(6e5), num_pc_evals // 2, base=10)
    return np.append(-np.flip(ep), ep)


class VStep:
    """Backward-Euler v-space step of the glow case (step_bte_v, ts_type "BE",
    glowdischarge_boltzmann_1d.py:1476-1592) on the device.

    C_en, C_gT, A_v: (dof, dof) device operators already premultiplied by M^-1 (the
    reference's op_col_en, op_col_gT, op_adv_v).  For a spatially varying field the
    system (I - dt tau (n0 np0 (C_en + Tg C_gT) + E(x) A_v)) y = u over all points is
    solved by right-preconditioned restarted GMRES with the per-point inverse taken
    from the table built at E-bin values (:2087-2112) and chosen by the nearest bin
    (:1429-1474); a uniform field is solved directly (:1561-1572)."""

    def __init__(self, C_en, C_gT, A_v, tau, n0, np0, Tg, dt, ordinates=None, Evals=None, rtol=1e-10, atol=1e-20,
                 gmres_rsrt=30):
        dev = C_en.device
        self.dev = dev
        self.dof = C_en.shape[0]
        self.tau, self.dt = float(tau), float(dt)
        self.K0 = (n0 * np0 * (C_en + Tg * C_gT)).contiguous()
        self.A = A_v.contiguous()
        self.ords = ordinates
        self.rtol, self.atol, self.rsrt = rtol, atol, int(gmres_rsrt)
        Ev = default_pc_fields() if Evals is None else np.asarray(Evals, dtype=np.float64)
        self.Evals = torch.from_numpy(Ev).to(dev)
        I = torch.eye(self.dof, dtype=torch.float64, device=dev)
        c = self.dt * self.tau
        mats = [I - c * Ev[0] * self.A - c * self.K0]
        for i in range(1, len(Ev)):
            mats.append(I - c * 0.5 * (Ev[i - 1] + Ev[i]) * self.A - c * self.K0)
        mats.append(I - c * Ev[-1] * self.A - c * self.K0)
        self.PmatE = torch.linalg.inv(torch.stack(mats)).contiguous()     # (len(Ev)+1, dof, dof), set-up only
        self.last_iters = 0
        self._gm = {}

    def _lmat(self, E):
        nx = E.shape[0]
        dof = self.dof
        c = self.dt * self.tau

        def mv(x):
            X = x.reshape(dof, nx)
            Kx = self.K0 @ X
            Ax = self.A @ X
            y = torch.empty_like(X)
            _lib.call("bz_vstep_combine", X.contiguous(), Kx, Ax, E, c, dof, nx, y)
            return y.reshape(-1)
        return mv

    def _pc(self, E):
        """Block preconditioner: each x point multiplies the inverse of its E bin's block
        (bz_nearest_bin picks the bin).  The bins are fixed for the whole v-step, so the points
        are grouped by bin once here; every application is then one batched FP64 GEMM over
        the bins in use (each block read once), between a gather and a scatter of the points."""
        nx = E.shape[0]
        bins = torch.empty(nx, dtype=torch.int32, device=self.dev)
        _lib.call("bz_nearest_bin", E, nx, self.Evals, self.Evals.shape[0], bins)
        dof = self.dof
        b64 = bins.long()
        order = torch.argsort(b64, stable=True)
        uniq, counts = torch.unique_consecutive(b64[order], return_counts=True)
        nb = int(uniq.numel())
        maxc = int(counts.max())
        start = torch.cumsum(counts, 0) - counts
        group = torch.repeat_interleave(torch.arange(nb, device=self.dev), counts, output_size=nx)
        pos = torch.arange(nx, device=self.dev) - start[group]
        idx = torch.zeros((nb, maxc), dtype=torch.long, device=self.dev)   # padding reads point 0
        idx[group, pos] = order
        Pg = self.PmatE[uniq].contiguous()                                  # (nb, dof, dof)

        def mv(x):
            X = x.reshape(dof, nx)
            U = X[:, idx].permute(1, 0, 2).contiguous()                     # (nb, dof, maxc)
            Yg = torch.bmm(Pg, U)                                           # (nb, dof, maxc)
            y = torch.empty((dof, nx), dtype=torch.float64, device=self.dev)
            y[:, order] = Yg[group, :, pos].t()
            return y.reshape(-1)
        return mv

    def solve_sph(self, u, E, use_graph=True):
        """u (dof, Nx) in the harmonic basis, E (Nx,) device -> y (dof, Nx).

        use_graph: the device GMRES (_DeviceGMRES: MFMA products, device residual estimate,
        restart cycles replayed from HIP graphs); otherwise krylov.gmres_pcr, the reference's
        loop with a host-checked true residual after every step."""
        from .krylov import gmres_pcr
        E = E.contiguous().to(torch.float64)
        nx = E.shape[0]
        if bool((E == E[0]).all()):
            L = torch.eye(self.dof, dtype=torch.float64, device=self.dev) - self.dt * self.tau * (self.K0 + E[0] * self.A)
            return torch.linalg.inv(L) @ u
        if use_graph:
            g = self._gm.get(nx)
            if g is None:
                g = self._gm[nx] = _DeviceGMRES(self, nx)
            y, iters = g.solve(u, E)
            self.last_iters = iters
            return y
        counter = [0]

        def cb(_):
            counter[0] += 1
        y, status = gmres_pcr(self._lmat(E), u.reshape(-1), x0=u.reshape(-1), atol=self.atol, rtol=self.rtol,
                              maxiter=self.rsrt * 50, restart=self.rsrt, M=self._pc(E), callback=cb)
        self.last_iters = counter[0]
        if status != 0:
            raise RuntimeError(f"GMRES did not converge in {status} iterations")
        return y.reshape(self.dof, nx)

    def step(self, v, E):
        """v (num_p*Nvt, Nx) ordinates -> ordinates after one BE v-step."""
        u = self.ords.ords_to_sph(v)
        return self.ords.sph_to_ords(self.solve_sph(u, E).contiguous())


class glow1d_vspace:
    """The velocity-space methods of glowdischarge_boltzmann_1d.glow1d_boltzmann with the
    reference's names and signatures, for a 1D3V time loop that keeps its own x-space and
    field solve: ords_to_sph / sph_to_ords (:1032-1048), initialize_bte_adv_x (:718-834),
    step_bte_x (:1220-1287) and step_bte_v (:1476-1592, ts_type "BE").

    op_col_en, op_col_gT, op_adv_v: the v-space operators premultiplied by M^-1, (dof_v,
    dof_v) device tensors, as the reference holds them; vgrid: the radial VGrid; xp: the Np
    spatial points on [-1, 1]; Nvt: ordinates; param: tau, n0, np0, Tg and L (gap length);
    args: rtol, atol and gmres_rsrt of the v-step GMRES.  The caller sets bs_E (Np,) on the
    device before step_bte_v, as the reference's loop does.  The v-step is the device GMRES
    (VStep); its E-bin preconditioner table is built for each new dt."""

    def __init__(self, op_col_en, op_col_gT, op_adv_v, vgrid, xp, Nvt, param, args):
        self.op_col_en, self.op_col_gT, self.op_adv_v = op_col_en, op_col_gT, op_adv_v
        self.vgrid = vgrid
        self.xp = np.asarray(xp, dtype=np.float64)
        self.Np, self.Nvt, self.Nr = len(self.xp), int(Nvt), vgrid.npr
        self.dof_v = op_col_en.shape[0]
        self.param, self.args = param, args
        self.ords = Ordinates(vgrid.npr, vgrid.lmax, self.Nvt, vgrid.dev)
        self.xp_vt, _ = ordinates(self.Nvt)
        self.adv_setup_dt = None
        self.bs_E = None
        self._xadv = None
        self._vsteps = {}

    def ords_to_sph(self, x):
        return self.ords.ords_to_sph(x)

    def sph_to_ords(self, x):
        return self.ords.sph_to_ords(x)

    def initialize_bte_adv_x(self, dt):
        """x-advection operators of a step of length dt (RF periods): the eigen basis of
        M^-1 (vth tau / (L/2)) [int v^3 B_p B_k] and the upwind backward-Euler factors."""
        scale = self.vgrid.vth * self.param.tau / (0.5 * self.param.L)
        self._xadv = XAdvection(self.vgrid, self.xp, self.Nvt, scale, dt)
        self.adv_setup_dt = dt

    def step_bte_x(self, v, time, dt, verbose=0):
        """v (Nr*Nvt, Np) ordinates -> advected over dt (the step the operators were set up for)."""
        assert self.adv_setup_dt == dt
        return self._xadv.step(v, dt)

    def step_bte_v(self, v, dv, time, dt, ts_type, verbose=0):
        """v (Nr*Nvt, Np) ordinates -> one backward-Euler v-step under the field bs_E."""
        if ts_type != "BE":
            raise NotImplementedError(f"ts_type {ts_type!r}: the v-step of this library is backward Euler")
        vs = self._vsteps.get(dt)
        if vs is None:
            p = self.param
            vs = VStep(self.op_col_en, self.op_col_gT, self.op_adv_v, p.tau, p.n0, p.np0, p.Tg, dt,
                       ordinates=self.ords, rtol=self.args.rtol, atol=self.args.atol,
                       gmres_rsrt=self.args.gmres_rsrt)
            self._vsteps[dt] = vs
        E = torch.as_tensor(self.bs_E, dtype=torch.float64, device=self.vgrid.dev).reshape(-1)
        return vs.step(v.contiguous(), E)
