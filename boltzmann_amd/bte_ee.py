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
the e-e term ne gamma_a(x) M^-1 [cc_a(x, H x) + cc_b(x, G x)] is bilinear in (x, potentials).
With the potentials and gamma_a frozen at the current iterate it is a banded linear operator
that is added to every instance's profile, and the fused Newton kernel takes a few steps on
the frozen problem; the outer iteration refreshes the potentials.  Its fixed point is the
reference's (same residual); convergence is tested on the full residual (potentials at the
iterate) against atol + rtol |r0| per instance.  The reference's Jacobian also differentiates
the potentials; the outer iteration replaces that term, so iteration counts differ.
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
        Minv = vgrid.mass_radial_inverse()
        self.H = torch.empty((ns, npr, npr), dtype=torch.float64, device=dev)
        self.G = torch.empty((ns, npr, npr), dtype=torch.float64, device=dev)
        raw = torch.empty((npr, npr), dtype=torch.float64, device=dev)
        for l in range(ns):
            m1 = torch.empty((npr, nq), dtype=torch.float64, device=dev)
            m2 = torch.empty_like(m1)
            _lib.call("bz_potential_kernels", l, float(m_ab), gxd, nq, npr, moment(l + 2, 0), moment(l + 4, 0),
                      moment(l - 1, 1), moment(l - 3, 1), m1, m2)
            for mk, dst in ((m1, self.H), (m2, self.G)):
                # raw[p, k] = sum_x w x^2 B_p(x) m(k, x);  H_l = M_r^-1 raw
                _lib.call("bz_gemm_strided", npr, npr, nq, 1.0, Vw, nq, 1, mk, 1, nq, raw, npr)
                _lib.call("bz_gemm_strided", npr, npr, npr, 1.0, Minv, npr, 1, raw, npr, 1, dst[l], npr)
        (ta_off, ta_c, ta_i), (tb_off, tb_c, tb_i) = _term_tables(vgrid.lmax, dev)
        self._tables = (ta_off, ta_c, ta_i, tb_off, tb_c, tb_i)
        glx, glw = vgrid.gl(vgrid.q)
        nc = ns ** 3
        self.ccA = torch.empty((npr, 7, 7, nc), dtype=torch.float64, device=dev)
        self.ccB = torch.empty_like(self.ccA)
        _lib.call("bz_cc_banded", vgrid.t, npr, ns, glx, glw, vgrid.q, ta_off, ta_c, ta_i, tb_off, tb_c, tb_i,
                  self.ccA, self.ccB)
        # Coulomb prefactor constants (bte_0d3v_batched.py:652-661)
        vth = vgrid.vth
        mw0 = 1.0 / (math.sqrt(math.pi) * vth) ** 3
        eps0, qe, me = sc.epsilon_0, sc.e, sc.electron_mass
        self.c_m0 = mw0 * vth ** 3
        self.c_kt = mw0 * vth ** 5 * sc.Boltzmann
        self.c_lam = 12.0 * math.pi * eps0 ** 1.5 / qe ** 3
        self.c_gam = qe ** 4 / (4.0 * math.pi * (eps0 * me) ** 2) / vth ** 3
        self.top = (solver.temp * TEMP_K_1EV).contiguous()
        self.anderson_depth = 8

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

    def steady_state(self, x, n0, Tg, E, ns, ne, atol, rtol, max_iter, alpha_min=1e-16, alpha_rho=0.5,
                     line_search_c=1e-4, inner_iter=1, max_outer=400, mode=1):
        """In place on x (n, B).  Returns a dict with per-instance iters (inner Newton steps),
        status (0 converged), res_abs (full residual at the end), res0 (full residual at the
        start) and the number of outer iterations."""
        B = x.shape[1]
        s = self.s
        tol = None
        r0 = None
        iters = torch.zeros(B, dtype=torch.int32, device=self.dev)
        outer = 0
        res = None
        max_outer = min(max_outer, max(int(max_iter), 1) + 1)
        # the potentials are those of the normalised iterate (u.x = 1), as in the reference
        x.div_((s.u @ x)[None, :])
        m_hist = self.anderson_depth
        n = self.n
        dX = torch.zeros((m_hist, n, B), dtype=torch.float64, device=self.dev)
        dF = torch.zeros_like(dX)
        x_prev = f_prev = None
        nh = 0
        while outer < max_outer:
            h, g, negam = self.potentials(x, ne)
            ee = _lib.EE0D(ns=self.ns, npr=self.npr, ccA=self.ccA.data_ptr(), ccB=self.ccB.data_ptr(),
                           h=h.data_ptr(), g=g.data_ptr(), negam=negam.data_ptr())
            x_old = x.clone()
            sc_ = s.solve_raw(x, n0, Tg, E, ns, atol=atol, rtol=rtol, max_iter=0 if tol is None else inner_iter,
                              alpha_min=alpha_min, alpha_rho=alpha_rho, ls_c=line_search_c, mode=mode, tol_abs=tol,
                              ee=ee)
            res = sc_["res0"].clone()               # full residual at the entry iterate
            outer += 1
            if tol is None:
                r0 = res
                tol = (atol + rtol * r0).contiguous()
                continue
            iters += sc_["iters"]
            if bool((res <= tol).all()):
                x.copy_(x_old)
                break
            # Anderson mixing of the fixed-point map x -> (frozen-potential Newton steps)(x)
            f = x - x_old
            if x_prev is not None:
                slot = nh % m_hist
                torch.sub(x_old, x_prev, out=dX[slot])
                torch.sub(f, f_prev, out=dF[slot])
                nh += 1
            x_prev, f_prev = x_old, f
            x.copy_(x_old)
            _lib.call("bz_anderson_mix", n, B, min(nh, m_hist), dX, dF, f.contiguous(), x)
        status = (res > tol).to(torch.int32)
        return {"iters": iters, "status": status, "res_abs": res, "res0": r0, "outer": outer}
