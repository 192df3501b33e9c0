"""Electron-electron collisions of the batched 0D solver on the MI355X.

Assembly (once per grid, HIP kernels of bz_ee.hip):
  * Rosenbluth potential operators H_l, G_l (the l blocks of hl_op, gl_op of
    CollisionOpSP.compute_rosenbluth_potentials_op, collision_operator_spherical.py:786-971),
    from the partial moments Pm / Qm at the quadrature points with the reference's Gauss
    orders, the potential kernels m1, m2 and the projection M^-1 [int v^2 B_p m(., k)];
  * the trilinear weak-form tensors cc_a, cc_b (coulomb_collision_op_assembly, :973-1050),
    kept banded ([npr][7][7][ns^3]): |k - p|, |r - p| <= 3 is where three cubic B-splines
    overlap, so nothing else is nonzero.  The dense e-e operator cc of the reference
    (n^3 doubles, 464 MB at n = 387, bte_0d3v_batched.py:462-468) is never formed.

Solve (bte_0d3v_batched.steady_state_solve with ee_collisions, :702-709, 734-743, 882-979):
Newton's method on F(x) = P (M^-1 K_x x - mu(x) x), K_x = K + ne gamma_a(x) [cc_a(., Hx) + cc_b(., Gx)],
with the reference's Jacobian (the potentials differentiated: cc x + cc_12 x; gamma_a held at the
iterate) plus the derivative of the growth term.  The dense Jacobian is never formed: each
Newton system is solved by GMRES, batched over the instances (one Krylov basis per instance),
right-preconditioned by the factorisation of the banded part K_x - mu M (the frozen-potential
operator, factorised per application with the fused solver's LU), and the potential-derivative
part T2 v = ne gamma_a [cc_a(x, Hv) + cc_b(x, Gv)] applied matrix-free.  Steps are damped by
Armijo backtracking on 0.5 |F|^2 per instance as in the reference; an instance stops at
|F| <= atol + rtol |F(x0)|.
"""
import math

import numpy as np
import scipy.constants as sc
import torch

from . import _lib
from .vspace import TEMP_K_1EV


def _term_tables(lmax, dev):
    from .ee_terms_table import LMAX, TERMS
    if lmax > LMAX:
        raise NotImplementedError(f"e-e term tables are generated up to lmax {LMAX}")
    ns = lmax + 1
    out = []
    for kind in ("a", "b"):
        off = [0]
        coef = []
        ints = []
        for q in range(ns):
            for l in range(ns):
                for s in range(ns):
                    for c, pw, dp, dk, dr in TERMS[kind].get((q, l, s), []):
                        coef.append(c)
                        ints.extend([pw, dp, dk, dr])
                    off.append(len(coef))
        out.append((torch.tensor(off, dtype=torch.int32, device=dev),
                    torch.tensor(coef if coef else [0.0], dtype=torch.float64, device=dev),
                    torch.tensor(ints if ints else [0, 0, 0, 0], dtype=torch.int32, device=dev)))
    return out


def rosenbluth_operators(vgrid, m_ab=1.0):
    """H_l, G_l (ns, npr, npr): the l blocks of hl_op, gl_op of
    CollisionOpSP.compute_rosenbluth_potentials_op (collision_operator_spherical.py:786-971),
    from the partial moments at the grid's quadrature points (the reference's Gauss orders),
    the potential kernels m1, m2 and the projection M^-1 [int v^2 B_p m(., k)], the last as
    a refined banded solve."""
    dev = vgrid.dev
    npr, ns = vgrid.npr, vgrid.num_sh
    # quadrature points of the grid: spline_qpts Gauss points per knot interval
    t = vgrid.t_np
    x, w = np.polynomial.legendre.leggauss(vgrid.q)
    gx, gw = [], []
    for i in range(len(t) - 1):
        if t[i] != t[i + 1]:
            gx.append(0.5 * (t[i + 1] - t[i]) * x + 0.5 * (t[i + 1] + t[i]))
            gw.append(0.5 * (t[i + 1] - t[i]) * w)
    gx = np.concatenate(gx)
    gw = np.concatenate(gw)
    nq = len(gx)
    gxd = torch.from_numpy(gx).to(dev)
    sp = vgrid.sp_order

    def moment(m, kind):
        qn = ((sp + abs(m) + 1) // 2 + 1) * (4 if kind == 0 else 32)
        qx, qw = np.polynomial.legendre.leggauss(qn)
        out = torch.empty((npr, nq), dtype=torch.float64, device=dev)
        _lib.call("bz_partial_moment", vgrid.t, npr, gxd, nq, int(m), kind, torch.from_numpy(qx).to(dev),
                  torch.from_numpy(qw).to(dev), qn, out)
        return out

    V = torch.empty((npr, nq), dtype=torch.float64, device=dev)
    _lib.call("bz_basis_table", vgrid.t, npr, gxd, nq, 0, V)
    Vw = (V * torch.from_numpy(gw * gx ** 2).to(dev)[None, :]).contiguous()
    mr = vgrid.mass_radial()
    lb = torch.empty((npr, 4), dtype=torch.float64, device=dev)
    _lib.call("bz_band_cholesky", mr, npr, 3, lb)
    work = torch.empty((2 * npr * npr,), dtype=torch.float64, device=dev)
    H = torch.empty((ns, npr, npr), dtype=torch.float64, device=dev)
    G = torch.empty((ns, npr, npr), dtype=torch.float64, device=dev)
    raw = torch.empty((npr, npr), dtype=torch.float64, device=dev)
    for l in range(ns):
        m1 = torch.empty((npr, nq), dtype=torch.float64, device=dev)
        m2 = torch.empty_like(m1)
        _lib.call("bz_potential_kernels", l, float(m_ab), gxd, nq, npr, moment(l + 2, 0), moment(l + 4, 0),
                  moment(l - 1, 1), moment(l - 3, 1), m1, m2)
        for mk, dst in ((m1, H), (m2, G)):
            # raw[p, k] = sum_x w x^2 B_p(x) m(k, x);  H_l = M_r^-1 raw as a refined solve
            # (the explicit inverse would carry cond(M_r) roundings into H, G)
            _lib.call("bz_gemm_strided", npr, npr, nq, 1.0, Vw, nq, 1, mk, 1, nq, raw, npr)
            _lib.call("bz_band_chol_solve_refined", mr, lb, npr, 3, raw, npr, dst[l], work, 3)
    return H, G


def cc_banded(vgrid):
    """The trilinear weak-form tensors cc_a, cc_b of coulomb_collision_op_assembly
    (collision_operator_spherical.py:973-1050), banded: (npr, 7, 7, ns^3) each, entry
    [p, k-p+3, r-p+3, (q ns + l) ns + s] of cc[(p,q), (k,l), (r,s)]."""
    npr, ns = vgrid.npr, vgrid.num_sh
    (ta_off, ta_c, ta_i), (tb_off, tb_c, tb_i) = _term_tables(vgrid.lmax, vgrid.dev)
    glx, glw = vgrid.gl(vgrid.q)
    nc = ns ** 3
    ccA = torch.empty((npr, 7, 7, nc), dtype=torch.float64, device=vgrid.dev)
    ccB = torch.empty_like(ccA)
    _lib.call("bz_cc_banded", vgrid.t, npr, ns, glx, glw, vgrid.q, ta_off, ta_c, ta_i, tb_off, tb_c, tb_i, ccA, ccB)
    return ccA, ccB


def cc_dense(vgrid, band):
    """Dense (n, n, n) tensor cc[(p,q), (k,l), (r,s)] from its banded storage (cc_banded)."""
    npr, ns, n = vgrid.npr, vgrid.num_sh, vgrid.n
    out = torch.zeros((npr, ns, npr, ns, npr, ns), dtype=torch.float64, device=vgrid.dev)
    b = band.reshape(npr, 7, 7, ns, ns, ns)
    p = torch.arange(npr, device=vgrid.dev)
    for dk in range(-3, 4):
        for dr in range(-3, 4):
            sel = (p + dk >= 0) & (p + dk < npr) & (p + dr >= 0) & (p + dr < npr)
            ps = p[sel]
            out[ps, :, ps + dk, :, ps + dr, :] = b[ps, dk + 3, dr + 3]
    return out.reshape(n, n, n)


class EEOperator:
    def __init__(self, vgrid, solver, m_ab=1.0):
        """vgrid: VGrid; solver: Batched0D built with ee_pattern=True (its mass / temperature
        vectors and solver tables are shared)."""
        self.g = vgrid
        self.s = solver
        dev = vgrid.dev
        self.dev = dev
        npr, ns, n = vgrid.npr, vgrid.num_sh, vgrid.n
        self.npr, self.ns, self.n = npr, ns, n
        self.H, self.G = rosenbluth_operators(vgrid, m_ab)
        self.ccA, self.ccB = cc_banded(vgrid)
        # Coulomb prefactor constants (bte_0d3v_batched.py:652-661)
        vth = vgrid.vth
        mw0 = 1.0 / (math.sqrt(math.pi) * vth) ** 3
        eps0, qe, me = sc.epsilon_0, sc.e, sc.electron_mass
        self.c_m0 = mw0 * vth ** 3
        self.c_kt = mw0 * vth ** 5 * sc.Boltzmann
        self.c_lam = 12.0 * math.pi * eps0 ** 1.5 / qe ** 3
        self.c_gam = qe ** 4 / (4.0 * math.pi * (eps0 * me) ** 2) / vth ** 3
        self.top = (solver.temp * TEMP_K_1EV).contiguous()
        self._ws = {}

    def potentials(self, x, ne):
        """h, g (B, n) and ne * gamma_a (B,) of the distributions x (n, B)."""
        B = x.shape[1]
        h = torch.empty((B, self.n), dtype=torch.float64, device=self.dev)
        g = torch.empty_like(h)
        negam = torch.empty(B, dtype=torch.float64, device=self.dev)
        _lib.call("bz_ee_potentials", self.H, self.G, self.ns, self.npr, x, B, h, g)
        _lib.call("bz_ee_gamma", x, self.n, B, self.s.mass, self.top, ne, self.c_m0, self.c_kt, self.c_lam,
                  self.c_gam, negam)
        return h, g, negam

    def potentials_dd(self, x, ne):
        """potentials() with the compensated sums: h, g, negam, h_lo, g_lo."""
        B = x.shape[1]
        h = torch.empty((B, self.n), dtype=torch.float64, device=self.dev)
        g, h_lo, g_lo = torch.empty_like(h), torch.empty_like(h), torch.empty_like(h)
        negam = torch.empty(B, dtype=torch.float64, device=self.dev)
        _lib.call("bz_ee_potentials_dd", self.H, self.G, self.ns, self.npr, x, B, h, g, h_lo, g_lo)
        _lib.call("bz_ee_gamma", x, self.n, B, self.s.mass, self.top, ne, self.c_m0, self.c_kt, self.c_lam,
                  self.c_gam, negam)
        return h, g, negam, h_lo, g_lo

    def residual_norm(self, x, prm, tg_scale=1.0):
        """|F(x)| (B,) of the Newton residual with e-e at x (n, B); prm = (n0, Tg, E, ns, ne)."""
        B = x.shape[1]
        h, g, negam, h_lo, g_lo = self.potentials_dd(x, prm[4])
        ee = _lib.EE0D(ns=self.ns, npr=self.npr, ccA=self.ccA.data_ptr(), ccB=self.ccB.data_ptr(), h=h.data_ptr(),
                       g=g.data_ptr(), negam=negam.data_ptr(), h_lo=h_lo.data_ptr(), g_lo=g_lo.data_ptr())
        bt, _ = self.s.batch_struct(x, *prm[:4], tg_scale=tg_scale)
        mu = torch.empty(B, dtype=torch.float64, device=self.dev)
        F = torch.empty((B, self.n), dtype=torch.float64, device=self.dev)
        fn = torch.empty_like(mu)
        _lib.call("bz_ee_newton_residual", self.s.tables.struct, bt, ee, mu, F, fn)
        return fn

    def steady_state(self, x, n0, Tg, E, ns, ne, atol, rtol, max_iter, alpha_min=1e-16, alpha_rho=0.5,
                     line_search_c=1e-4, restart=30, forcing=1e-10, tg_scale=1.0, callback=None,
                     floor_rel=1e-8, floor_alpha=1.0 / 64):
        """Newton-Krylov on every instance, in place on x (n, B) (normalised to u.x = 1 first,
        as the reference does).  Stops an instance at |F| <= atol + rtol |F(x0)|.  Returns a
        dict with per-instance iters (Newton steps), krylov (GMRES steps), status (0 converged,
        1 max_iter, 2 line search failed, 3 stopped at the FP64 floor), res_abs, res0 and the
        number of Newton sweeps.

        The FP64 evaluation floor of |F| is absolute (rounding in K_x x and the mass solve),
        while r0 scales with the e-e strength, so atol + rtol r0 can lie below it (measured at
        n = 50: 1e-12 r0 to 8e-12 r0).  An instance whose |F| is below floor_rel r0 and whose
        line search finds no decrease down to a step of floor_alpha stops there (status 3);
        the reference would iterate on to max_iter at that point.
        callback(sweep, instance indices, |F|, step lengths, Krylov steps) over the working set
        after every sweep.

        forcing: each Newton system is solved to |J d + F| <= max(forcing |F|, tol / 4).  The
        Jacobian is ill-conditioned (measured at n = 50: a linear residual of 1e-4 |F| leaves
        15-80 % error in the step), so the default asks for a nearly exact step; 12-14 Krylov
        steps reach it with the frozen-potential preconditioner."""
        B = x.shape[1]
        s = self.s
        n, dev, f64, i32 = self.n, self.dev, torch.float64, torch.int32
        m = int(restart)
        x.div_((s.u @ x)[None, :])
        grid = s.tables.struct
        iters = torch.zeros(B, dtype=i32, device=dev)
        krylov = torch.zeros(B, dtype=i32, device=dev)
        failed = torch.zeros(B, dtype=torch.bool, device=dev)
        floor = torch.zeros(B, dtype=torch.bool, device=dev)
        res = torch.empty(B, dtype=f64, device=dev)

        # Working set: the instances still iterating, gathered into a dense sub-batch whenever
        # fewer than half of the current one remain, so that the slow tail of a batch does not
        # pay for whole-batch sweeps.
        idx = torch.arange(B, device=dev)
        xw, prm = x, (n0, Tg, E, ns, ne)
        w = self._ws.get((B, m))
        if w is None:
            self._ws = {(B, m): _NKWork(n, B, m, dev, s.tables.P)}
            w = self._ws[(B, m)]

        def evaluate(xe, w, prm):
            """potentials, profiles K_x - mu M, F and |F| at xe; returns the batch struct and negam"""
            h, g, negam, h_lo, g_lo = self.potentials_dd(xe, prm[4])
            ee = _lib.EE0D(ns=self.ns, npr=self.npr, ccA=self.ccA.data_ptr(), ccB=self.ccB.data_ptr(),
                           h=h.data_ptr(), g=g.data_ptr(), negam=negam.data_ptr(), h_lo=h_lo.data_ptr(),
                           g_lo=g_lo.data_ptr())
            bt, _ = s.batch_struct(xe, *prm[:4], tg_scale=tg_scale)
            _lib.call("bz_ee_newton_residual", grid, bt, ee, w.mu, w.F, w.fn)
            return bt, negam, (h, g, h_lo, g_lo)

        bt, negam, keep = evaluate(xw, w, prm)
        r0 = w.fn.clone()
        tol = atol + rtol * r0
        tolw = tol
        x_try = torch.empty_like(xw)
        sweeps = 0
        while True:
            Bw = xw.shape[1]
            todo = (w.fn > tolw) & ~failed[idx] & ~floor[idx]
            if sweeps >= max_iter:
                break
            nt = int(todo.sum())
            if nt == 0:
                break
            if nt <= Bw // 2 and Bw > 64:
                res[idx] = w.fn
                if xw is not x:
                    x[:, idx] = xw
                idx = idx[todo]
                xw = x[:, idx].contiguous()
                prm = tuple(v[..., idx].contiguous() for v in (n0, Tg, E, ns, ne))
                tolw = tol[idx].contiguous()
                w = _NKWork(n, nt, m, dev, s.tables.P)
                x_try = torch.empty_like(xw)
                bt, negam, keep = evaluate(xw, w, prm)
                Bw = nt
                todo = (w.fn > tolw) & ~failed[idx] & ~floor[idx]
            sweeps += 1
            iters.index_add_(0, idx, todo.to(i32))
            # ---- GMRES on J d = -F, right-preconditioned by the frozen-potential factorisation
            w.active.copy_(todo.to(i32))
            _lib.call("bz_ee_newton_factor", grid, bt, w.active, w.lprof, w.z1, w.s1)
            w.kdim.zero_()
            w.g.zero_()
            w.g[:, 0] = w.fn
            torch.maximum(forcing * w.fn, 0.25 * tolw, out=w.target)
            torch.div(w.F, -w.fn.clamp_min(1e-300)[:, None], out=w.V[0])
            for i in range(m):
                _lib.call("bz_ee_newton_precond_apply", grid, bt, w.active, w.lprof, w.z1, w.s1, w.V[i], w.z, w.zT)
                _lib.call("bz_ee_potentials_cols", self.H, self.G, self.ns, self.npr, w.zT, Bw, w.hz, w.gz)
                _lib.call("bz_ee_jacobian_t2", self.ns, self.npr, Bw, self.ccA, self.ccB, xw, w.hz, w.gz, negam,
                          w.active, w.t2)
                _lib.call("bz_ee_newton_jv", grid, bt, w.active, w.z, w.t2, w.Jz)
                _lib.call("bz_arnoldi_batched", n, Bw, m, i, w.V, w.Jz, w.R, w.cs, w.sn, w.g, w.target, w.active,
                          w.kdim, w.resid)
                if not bool(w.active.any()):
                    break
            krylov.index_add_(0, idx, w.kdim)
            _lib.call("bz_gmres_combine_batched", n, Bw, m, w.V, w.R, w.g, w.kdim, w.Jz)
            w.active.copy_((w.kdim > 0).to(i32))
            w.z.zero_()
            w.zT.zero_()
            _lib.call("bz_ee_newton_precond_apply", grid, bt, w.active, w.lprof, w.z1, w.s1, w.Jz, w.z, w.zT)
            # ---- Armijo backtracking on 0.5 |F|^2 per instance (bte_0d3v_batched.py:936-952)
            fn0 = w.fn.clone()
            r0w = r0[idx]
            alpha = torch.ones(Bw, dtype=f64, device=dev)
            pending = todo.clone()
            stopped = torch.zeros(Bw, dtype=torch.bool, device=dev)
            accepted_first = True
            # first trial (alpha = 1) over the whole working set: when every instance takes it,
            # its profiles and residuals are those of the next sweep
            torch.addcmul(xw, w.zT, alpha[None, :], out=x_try)
            bt, negam, keep = evaluate(x_try, w, prm)
            ok = w.fn * w.fn <= fn0 * fn0 * (1.0 - 2.0 * line_search_c * alpha)
            torch.where((pending & ok)[None, :], x_try, xw, out=xw)
            pending &= ~ok
            while bool(pending.any()):
                accepted_first = False
                alpha = torch.where(pending, alpha * alpha_rho, alpha)
                flat = pending & (alpha < floor_alpha) & (fn0 <= floor_rel * r0w)
                dead = pending & (alpha < alpha_min)
                floor.index_copy_(0, idx, floor[idx] | flat)
                failed.index_copy_(0, idx, failed[idx] | dead)
                stopped |= flat | dead
                pending &= ~(dead | flat)
                pid = pending.nonzero().squeeze(1)
                if pid.numel() == 0:
                    break
                # later trials only over the instances still backtracking
                xs = torch.addcmul(xw[:, pid], w.zT[:, pid], alpha[pid][None, :]).contiguous()
                fs = self.residual_norm(xs, tuple(v[..., pid].contiguous() for v in prm), tg_scale)
                oks = fs * fs <= fn0[pid] * fn0[pid] * (1.0 - 2.0 * line_search_c * alpha[pid])
                tk = pid[oks]
                xw[:, tk] = xs[:, oks]
                pending[tk] = False
            if callback is not None:
                callback(sweeps, idx, w.fn.clone(), alpha, w.kdim.clone())
            if not accepted_first or bool(stopped.any()):
                bt, negam, keep = evaluate(xw, w, prm)
            else:
                xw.copy_(x_try)
                bt, _ = s.batch_struct(xw, *prm[:4], tg_scale=tg_scale)
        res[idx] = w.fn
        if xw is not x:
            x[:, idx] = xw
        status = torch.where(res <= tol, 0, torch.where(failed, 2, torch.where(floor, 3, 1))).to(i32)
        return {"iters": iters, "krylov": krylov, "status": status, "res_abs": res, "res0": r0, "outer": sweeps}


class _NKWork:
    """Device work arrays of the batched Newton-Krylov solve (B instances of size n, restart m)."""

    def __init__(self, n, B, m, dev, P=1):
        f = dict(dtype=torch.float64, device=dev)
        self.lprof = torch.empty((B, P), **f)        # L of the preconditioner, profile positions
        self.z1 = torch.empty((B, n), **f)
        self.s1 = torch.empty(B, **f)
        self.mu = torch.empty(B, **f)
        self.F = torch.empty((B, n), **f)
        self.fn = torch.empty(B, **f)
        self.V = torch.empty((m + 1, B, n), **f)
        self.R = torch.zeros((B, m, m), **f)
        self.cs = torch.zeros((B, m), **f)
        self.sn = torch.zeros((B, m), **f)
        self.g = torch.zeros((B, m + 1), **f)
        self.target = torch.empty(B, **f)
        self.resid = torch.empty(B, **f)
        self.active = torch.zeros(B, dtype=torch.int32, device=dev)
        self.kdim = torch.zeros(B, dtype=torch.int32, device=dev)
        self.z = torch.empty((B, n), **f)
        self.zT = torch.empty((n, B), **f)
        self.hz = torch.empty((n, B), **f)
        self.gz = torch.empty((n, B), **f)
        self.t2 = torch.empty((B, n), **f)
        self.Jz = torch.empty((B, n), **f)
