"""Oracle: the fully implicit 1D3V Boltzmann step on the finite-volume / ordinate velocity
grid, with the XVTVRC operator-split preconditioner (FP64, numpy, dense matrices).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates (BESolver/python/...):
  * fd_coefficients / upwinded_dx / upwinded_dvt        mesh.py:11-56, 80-106
    (the reference differentiates scipy Lagrange polynomials; here the same nodal
    derivative weights come from a centred, scaled Vandermonde solve)
  * the radial grid: q Gauss points per knot interval   spec_spherical.py:1010-1012,
                                                        basis.py:414-429
  * the ordinates, "npsp" half-range Gauss points        spec_spherical.py:1060-1064
  * the strong-form collision operator per process      collision_operator_spherical.py:1334-1419
  * the upwinded (vr, vt) E-field advection Ep / En     spec_spherical.py:876-954
  * operator scalings, QoI vectors                      bte_1d3v_solver.py:440-505
  * the x shift operator, BE_UPW_FD                     bte_1d3v_solver.py:780-862
  * init: PmatC and the band tables                     bte_1d3v_solver.py:1333-1502
  * step_op_split, XVTVRC                               bte_1d3v_solver.py:1745-1870
  * step, FULLY_IMPLICIT: rhs_mvec and GMRES            bte_1d3v_solver.py:1872-1974
  * the Maxwellian initial state, compute_qoi,
    normalize_edf                                       bte_1d3v_solver.py:1179-1216,
                                                        1296-1326, 2055-2061
The state is v (Nr*Nvt, Nx): radial point r, ordinate t, point x, x fastest.

Pinned by tests/test_oracle_fvm1d3v.py: the difference weights are exact on polynomials
of the stencil's degree; the advection operator applied to g(v cos(theta)) reproduces
g'(v cos(theta)) (the derivative along the field) away from its boundary rows, at the
stencil's order; the collision operator conserves particles for elastic and excitation
and creates one electron per ionization, against the rate vector, to discretization
accuracy; the x shift operator is I + dt a d/dx exactly on linear functions; the XVTVRC
preconditioner is the exact inverse when a single operator is active; the GMRES step
solves the dense system.
"""
import math

import numpy as np
import scipy.constants as sc

from . import bspline
from .krylov import gmres_pcr
from .operators import C_GAMMA, QE, ME, KB, TEMP_K_1EV, vth_from_Te

QM = QE / ME


# ----------------------------------------------------------------------------- difference weights

def fd_coefficients(x, order):
    """mesh.py:11-17: W[i, j] = d^order L_j / dx^order at x_i, L_j the Lagrange polynomial of
    node j on the points x."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    c = x.mean()
    h = max(np.abs(x - c).max(), 1e-300)
    z = (x - c) / h
    V = np.vander(z, n, increasing=True)                   # V[i, k] = z_i^k
    Vd = np.zeros((n, n))
    for k in range(order, n):
        Vd[:, k] = math.factorial(k) / math.factorial(k - order) * z ** (k - order)
    # p(z) = sum_k a_k z^k with V a = f  =>  p^(m)(z_i) = Vd V^-1 f
    return np.linalg.solve(V.T, Vd.T).T / h ** order


def upwinded_dx(x, dorder, sw, dir):
    """mesh.py:20-56: one-sided sw-point stencils; "L" uses the points to the left of (and
    at) row i, "R" those to the right, with shorter stencils at the start/end."""
    Np = len(x)
    Dx = np.zeros((Np, Np))
    if dir == "L":
        for i in range(1, sw):
            Dx[i, 0:i + 1] = fd_coefficients(x[0:i + 1], dorder)[-1]
        for i in range(sw, Np):
            Dx[i, i - (sw - 1):i + 1] = fd_coefficients(x[i - (sw - 1):i + 1], dorder)[-1]
    else:
        assert dir == "R"
        for i in range(0, Np - (sw - 1)):
            Dx[i, i:i + sw] = fd_coefficients(x[i:i + sw], dorder)[0]
        for i in range(1, sw):
            Dx[-(i + 1), -(i + 1):] = fd_coefficients(x[-(i + 1):], dorder)[0]
    return Dx


def upwinded_dvt(x, dorder, sw, dir):
    """mesh.py:80-106 with use_cdx_internal=False: upwinded_dx on each half of the ordinates
    and a two-point row joining the halves (row mp for "L", row mp-1 for "R")."""
    Np = len(x)
    assert Np % 2 == 0
    mp = Np // 2
    Dx = np.zeros((Np, Np))
    Dx[0:mp, 0:mp] = upwinded_dx(x[0:mp], 1, sw, dir)
    Dx[mp:, mp:] = upwinded_dx(x[mp:], 1, sw, dir)
    if dir == "L":
        Dx[mp] = 0.0
        Dx[mp, mp - 1:mp + 1] = fd_coefficients(x[mp - 1:mp + 1], 1)[-1]
    else:
        Dx[mp - 1] = 0.0
        Dx[mp - 1, mp - 1:mp + 1] = fd_coefficients(x[mp - 1:mp + 1], 1)[0]
    return Dx


# ----------------------------------------------------------------------------- grids

def vt_grid(Nvt):
    """spec_spherical.py:1060-1064 (mode "npsp"): arccos of Nvt/2 Gauss points on each half
    of [-1, 1]; theta descends, cos(theta) < 0 on the first half."""
    gx, gw = np.polynomial.legendre.leggauss(Nvt // 2)
    mu = np.append(0.5 * gx - 0.5, 0.5 * gx + 0.5)
    return np.arccos(mu), np.append(0.5 * gw, 0.5 * gw)


def vr_grid(t, q):
    """spec_spherical.py:1010-1012 with use_bspline_qgrid: q Gauss points per knot interval."""
    return bspline.gauss_pn(t, q)


def cheb_points(Np):
    """bte_1d3v_solver.py:253-256 (adaptive grid): -cos(pi k / (Np-1))."""
    return -np.cos(np.pi * np.linspace(0, Np - 1, Np) / (Np - 1))


# ----------------------------------------------------------------------------- operators

def collision_fvm(vr, vr_qw, vt_qw, proc, vth, domain):
    """collision_operator_spherical.py:1334-1419 for one process.  Returns (G, dcs, C):
    the gain matrix G = diag(dcs(v_pre) detJ) Lvr (Nr x Nr), the loss coefficients dcs
    (the operator's loss diagonal is 4 pi dcs) and the dense operator
    C = kron(G, f0_op) - 4 pi kron(diag(dcs), I), f0_op[t, t'] = 2 pi w_t'."""
    Nr, Nvt = len(vr), len(vt_qw)
    rmax = np.sum(vr_qw)
    assert abs(1 - rmax / (domain[1] - domain[0])) < 1e-14
    th = proc.threshold
    if proc.type == "ELASTIC":
        v_scale = math.sqrt(1 - 2 * proc.mByM)
        vr_pre = vr / v_scale
        detJ = (1 / v_scale) ** 3
    elif proc.type == "EXCITATION":
        vr_pre = np.sqrt(vr ** 2 + (C_GAMMA / vth) ** 2 * th)
        detJ = vr_pre / vr
    elif proc.type == "IONIZATION":
        vr_pre = np.sqrt(2 * vr ** 2 + (C_GAMMA / vth) ** 2 * th)
        detJ = 4.0 * vr_pre / vr
    else:
        raise NotImplementedError(proc.type)
    dcs = (1 / (4 * np.pi)) * vr * vth * proc.total_cross_section((vr * vth / C_GAMMA) ** 2)
    dcs_pre = (1 / (4 * np.pi)) * vr_pre * vth * proc.total_cross_section((vr_pre * vth / C_GAMMA) ** 2)
    # linear interpolation of the grid functions at v_pre, zero outside [vr[0], vr[-1]]
    Lvr = np.zeros((Nr, Nr))
    for i, p in enumerate(vr_pre):
        if p < vr[0] or p > vr[-1]:
            continue
        k = int(np.searchsorted(vr, p, side="left"))
        k = min(max(k, 1), Nr - 1)
        lo, hi = k - 1, k
        s = (p - vr[lo]) / (vr[hi] - vr[lo])
        Lvr[i, lo] = 1.0 - s
        Lvr[i, hi] = s
    bidx = np.arange(Nr)[np.logical_and(vr_pre > vr[-1], vr_pre <= rmax)]
    Lvr[bidx, -1] = (rmax - vr_pre[bidx]) / (rmax - vr[-1])
    G = (dcs_pre * detJ)[:, None] * Lvr
    f0_op = np.tile(2 * np.pi * vt_qw, (Nvt, 1))
    C = np.kron(G, f0_op) - 4 * np.pi * np.kron(np.diag(dcs), np.eye(Nvt))
    return G, dcs, C


def vr_derivatives(vr, sw, domain_end):
    """The radial one-sided derivatives with the reference's boundary rows
    (spec_spherical.py:910-918, bte_1d3v_solver.py:1412-1421): zero-gradient inflow at the
    first point (left to right) and zero inflow beyond the last (right to left)."""
    DL = upwinded_dx(vr, 1, sw, "L")
    DR = upwinded_dx(vr, 1, sw, "R")
    DL[0, :] = 0.0
    DR[-1, :] = 0.0
    DR[-1, -1] = -1 / (domain_end - vr[-1])
    return DL, DR


def vt_derivatives(vt, sw):
    """Regularity at the poles: first row of the left-to-right and last row of the right-to-
    left angular derivative are zero (spec_spherical.py:922-923)."""
    DL = upwinded_dvt(vt, 1, sw, "L")
    DR = upwinded_dvt(vt, 1, sw, "R")
    DL[0, :] = 0.0
    DR[-1, :] = 0.0
    return DL, DR


def advection_fvm_parts(vr, vt, sw_vr, sw_vt, domain):
    """The radial and angular parts of advection_fvm: (Ep_vr, Ep_vt, En_vr, En_vt) with
    Ep = Ep_vr - Ep_vt, En = En_vr - En_vt."""
    Nr, Nvt = len(vr), len(vt)
    cos, sin = np.cos(vt), np.sin(vt)
    DvtL, DvtR = vt_derivatives(vt, sw_vt)
    DvrL, DvrR = vr_derivatives(vr, sw_vr, domain[1])
    mp = Nvt // 2
    assert (cos[mp:] > 0).all() and (cos[:mp] < 0).all()
    Ep = np.zeros((Nr, Nvt, Nr, Nvt))
    En = np.zeros((Nr, Nvt, Nr, Nvt))
    Ep[:, mp:, :, mp:] = np.kron(DvrR, np.diag(cos[mp:])).reshape((Nr, mp, Nr, mp))
    Ep[:, :mp, :, :mp] = np.kron(DvrL, np.diag(cos[:mp])).reshape((Nr, mp, Nr, mp))
    En[:, mp:, :, mp:] = np.kron(DvrL, np.diag(cos[mp:])).reshape((Nr, mp, Nr, mp))
    En[:, :mp, :, :mp] = np.kron(DvrR, np.diag(cos[:mp])).reshape((Nr, mp, Nr, mp))
    n = Nr * Nvt
    return (Ep.reshape((n, n)), np.kron(np.diag(1 / vr), np.diag(sin) @ DvtR),
            En.reshape((n, n)), np.kron(np.diag(1 / vr), np.diag(sin) @ DvtL))


def advection_fvm(vr, vt, sw_vr, sw_vt, domain):
    """spec_spherical.py:876-954 (use_upwinding=True): (Ep, En), the upwinded discretisations
    of cos(theta) d/dvr - (sin(theta)/vr) d/dtheta for E >= 0 and E < 0."""
    Ep_vr, Ep_vt, En_vr, En_vt = advection_fvm_parts(vr, vt, sw_vr, sw_vt, domain)
    return Ep_vr - Ep_vt, En_vr - En_vt


def x_derivatives(xp, px):
    """bte_1d3v_solver.py:801-817 (BE_UPW_FD): one-sided derivatives with zero inflow rows
    (first row of DpL, last row of DpR) and shorter stencils next to the boundary."""
    Nx = len(xp)
    DpL = np.zeros((Nx, Nx))
    DpR = np.zeros((Nx, Nx))
    DpL[1:, :] = upwinded_dx(xp, 1, px + 1, "L")[1:, :]
    DpR[:-1, :] = upwinded_dx(xp, 1, px + 1, "R")[:-1, :]
    if px > 1:
        for i in range(1, px + 1):
            DpL[i, :] = 0.0
            DpR[-(i + 1), :] = 0.0
            DpL[i, 0:i + 1] = fd_coefficients(xp[0:i + 1], 1)[-1]
            DpR[-(i + 1), -(i + 1):] = fd_coefficients(xp[-(i + 1):], 1)[0]
    return DpL, DpR


def x_shift_diag(DpL, DpR, adx, vt, dt, px):
    """bte_1d3v_solver.py:844-860: the bands of I + dt adx_r cos(theta_t) Dp, Dp = DpL for
    theta <= pi/2 (lower bands, [k, i] couples i to i-k) else DpR (upper, i to i+k).
    Returns (Nr, Nvt, px+1, Nx)."""
    Nr, Nvt, Nx = len(adx), len(vt), DpL.shape[0]
    out = np.zeros((Nr, Nvt, px + 1, Nx))
    for j in range(Nvt):
        low = vt[j] <= 0.5 * np.pi
        D = DpL if low else DpR
        for i in range(Nr):
            c = dt * adx[i] * np.cos(vt[j])
            out[i, j, 0] = np.ones(Nx) + c * np.diagonal(D, offset=0)
            for k in range(1, px + 1):
                if low:
                    out[i, j, k, k:] = c * np.diagonal(D, offset=-k)
                else:
                    out[i, j, k, :-k] = c * np.diagonal(D, offset=k)
    return out


def column_bands(D, nd, lower):
    """The reference's band tables of the sweep kernels (bte_1d3v_solver.py:1440-1450):
    lower: B[k, j] = D[j+k, j];  upper: B[k, j] = D[j-k, j]."""
    N = D.shape[0]
    B = np.zeros((nd, N))
    for k in range(nd):
        if lower:
            B[k, 0:N - k] = np.diagonal(D, offset=-k)
        else:
            B[k, k:] = np.diagonal(D, offset=k)
    return B


# ----------------------------------------------------------------------------- the solver

class Params:
    """The [bte] parameters the step reads (bte_1d3v_solver.py:77-216)."""

    def __init__(self, L=2.54e-2, V0=100.0, freq=13.56e6, Tg=300.0, p0=1.0, Np=16, Nr=15, Nvt=8, sp_order=3,
                 spline_qpts=3, ev_max=80.0, ev_extend=0, Te=4.0, dt=1e-3, pvr=3, pvt=3, px=1, atol=1e-12,
                 rtol=1e-10, gmres_rsrt=5, gmres_rsrt_iter=10):
        self.L = 0.5 * L
        self.V0, self.f, self.tau = V0, freq, 1.0 / freq
        self.qe, self.me, self.kB = QE, ME, KB
        self.Tg, self.p0 = Tg, p0
        self.n0 = (p0 * sc.torr / (KB * Tg) if Tg != 0 else 3.22e22 * p0)
        self.np0 = 8e16
        self.n0 = self.n0 / self.np0
        self.Np, self.Nr, self.Nvt = Np, Nr, Nvt
        self.sp_order, self.spline_qpts = sp_order, spline_qpts
        self.ev_max, self.ev_extend, self.Te, self.dt = ev_max, ev_extend, Te, dt
        self.pvr, self.pvt, self.px = pvr, pvt, px
        self.atol, self.rtol = atol, rtol
        self.gmres_rsrt, self.gmres_maxiter = gmres_rsrt, gmres_rsrt_iter * gmres_rsrt


class Fvm1d3v:
    """bte_1d3v with vspace FVM, solver FULLY_IMPLICIT, pc XVTVRC, xspace BE_UPW_FD."""

    def __init__(self, prm: Params, procs):
        self.params = p = prm
        self.procs = procs
        self.vth = vth = vth_from_Te(p.Te)
        self.xp = cheb_points(p.Np)
        kd = (0.0, math.sqrt(p.ev_max) * C_GAMMA / self.vth)
        t, domain = bspline.make_knots(kd, p.sp_order, p.Nr + 1, None, extend_domain=(p.ev_extend == 1),
                                       extend_domain_with_log=(p.ev_extend == 2))
        self.kdomain, self.domain = kd, domain
        self.vr, self.vr_qw = vr_grid(t, p.spline_qpts)
        self.vt, self.vt_qw = vt_grid(p.Nvt)
        assert self.vr[0] > domain[0] and self.vr[-1] < domain[1]
        Nr, Nvt, Nx = len(self.vr), p.Nvt, p.Np
        self.Nr, self.Nvt, self.Nx = Nr, Nvt, Nx
        self.dof_v = Nr * Nvt
        FOp = 0.0
        self.G = np.zeros((Nr, Nr))
        self.loss = np.zeros(Nr)
        for pr in procs:
            G, dcs, C = collision_fvm(self.vr, self.vr_qw, self.vt_qw, pr, vth, domain)
            FOp = FOp + C
            self.G += G
            self.loss += 4 * np.pi * dcs
        self.op_col_en = FOp
        Ep, En = advection_fvm(self.vr, self.vt, p.pvr + 1, p.pvt + 1, domain)
        self.op_adv_v_Ep = (1 / vth) * QM * Ep
        self.op_adv_v_En = (1 / vth) * QM * En
        self.op_adv_x_d = self.vr * vth * (p.tau / p.L)
        ev_fac = 0.5 * ME * (2 / 3 / KB) / TEMP_K_1EV
        self.op_mass = 2 * np.pi * np.kron(self.vr ** 2 * self.vr_qw, self.vt_qw)
        self.op_temp = 2 * np.pi * vth ** 2 * np.kron(self.vr ** 4 * self.vr_qw, self.vt_qw) * ev_fac
        self.op_rate = [(2 * np.pi) * (2 * vth ** 4 / C_GAMMA ** 3) *
                        np.kron(self.vr ** 3 * self.vr_qw * pr.total_cross_section((self.vr * vth / C_GAMMA) ** 2),
                                self.vt_qw) for pr in procs]

    # -------------------------------------------------------------- state and QoIs
    def initial_condition(self):
        """bte_1d3v_solver.py:1277-1326 (type 0, no IC file)."""
        p = self.params
        xx = p.L * (self.xp + 1)
        ne = 1e6 * (1e7 + 1e9 * (1 - 0.5 * xx / p.L) ** 2 * (0.5 * xx / p.L) ** 2) / p.np0
        v = np.zeros((self.dof_v, self.Nx))
        v_ratio = C_GAMMA * math.sqrt(p.Te) / self.vth
        hv = (1 / np.sqrt(np.pi) ** 3) * np.exp(-((self.vr / v_ratio) ** 2)) / v_ratio ** 3
        for i in range(self.Nx):
            v[:, i] = np.kron(hv, np.ones(self.Nvt))
        return ne[None, :] * v

    def normalize_edf(self, v):
        return v / (self.op_mass @ v * (2 * (self.vth / C_GAMMA) ** 3))

    def compute_qoi(self, v):
        """bte_1d3v_solver.py:1179-1216 (FVM branch)."""
        m0 = self.op_mass @ v
        Te = (self.op_temp @ v) / m0
        vn = self.normalize_edf(v)
        rates = np.array([r @ vn for r in self.op_rate]).T
        return {"rates": rates, "ne": m0 * self.params.np0, "Te": Te}

    # -------------------------------------------------------------- init
    def init(self, dt, gmres_pc_cv=1.0, gmres_pc_ax=1.0):
        """bte_1d3v_solver.py:1333-1502 for FVM: PmatC = (I - pc_cv dt tau n0 np0 C)^-1 (the
        dense inverse the reference takes for dof_v <= 65536), the x shift bands at
        pc_ax dt, and the band tables of the sweeps."""
        p = self.params
        vmat = p.tau * p.n0 * p.np0 * self.op_col_en
        self.PmatC = np.linalg.inv(np.eye(self.dof_v) - gmres_pc_cv * dt * vmat)
        DpL, DpR = x_derivatives(self.xp, p.px)
        self.bte_x_shift_diag = x_shift_diag(DpL, DpR, self.op_adv_x_d, self.vt, gmres_pc_ax * dt, p.px)
        DvtL, DvtR = vt_derivatives(self.vt, p.pvt + 1)
        DvrL, DvrR = vr_derivatives(self.vr, p.pvr + 1, self.kdomain[1])
        DxL = upwinded_dx(self.xp, 1, p.px + 1, "L")
        DxR = upwinded_dx(self.xp, 1, p.px + 1, "R")
        DxL[0, :] = 0.0
        DxR[-1, :] = 0.0
        sin = np.sin(self.vt)
        self.vr_diag_l2r = column_bands(DvrL, p.pvr + 1, True)
        self.vr_diag_r2l = column_bands(DvrR, p.pvr + 1, False)
        self.vt_diag_l2r = column_bands(np.diag(sin) @ DvtL, p.pvt + 1, True)
        self.vt_diag_r2l = column_bands(np.diag(sin) @ DvtR, p.pvt + 1, False)
        self.x_diag_l2r = column_bands(DxL, p.px + 1, True)
        self.x_diag_r2l = column_bands(DxR, p.px + 1, False)
        self.idx_np_vt = self.vt <= 0.5 * np.pi
        self.idx_sp_vt = self.vt > 0.5 * np.pi

    # -------------------------------------------------------------- operators of the step
    def step_op_split(self, E, v, dt, alpha=1.0):
        """bte_1d3v_solver.py:1745-1870, XVTVRC: collisions (PmatC), then backward-Euler sweeps
        in vr, vt and x (inflow values zeroed before the x sweep)."""
        from .banded import adv_vr_BE, adv_vt_BE, adv_x_BE
        p = self.params
        Nr, Nvt, Nx = self.Nr, self.Nvt, self.Nx
        aE = (p.tau * p.qe / p.me / self.vth) * E
        ax = self.vth * (p.tau / p.L)
        yc = self.PmatC @ v.reshape(self.dof_v, Nx)
        yc = np.swapaxes(yc.reshape(Nr, Nvt, Nx), 0, 2).reshape(-1)           # (Nx, Nvt, Nr)
        yvr = adv_vr_BE(self.vr_diag_l2r, self.vr_diag_r2l, aE, np.cos(self.vt), alpha * dt, yc, Nr, Nvt, Nx)
        yvr = np.swapaxes(yvr.reshape(Nx, Nvt, Nr), 1, 2).reshape(-1)          # (Nx, Nr, Nvt)
        yvt = adv_vt_BE(self.vt_diag_l2r, self.vt_diag_r2l, aE, self.vr, alpha * dt, yvr, Nr, Nvt, Nx)
        yvt = np.swapaxes(np.swapaxes(yvt.reshape(Nx, Nr, Nvt), 0, 2), 0, 1).copy()   # (Nr, Nvt, Nx)
        yvt[:, self.idx_np_vt, 0] = 0.0
        yvt[:, self.idx_sp_vt, -1] = 0.0
        yx = adv_x_BE(self.x_diag_l2r, self.x_diag_r2l, ax * self.vr, np.cos(self.vt), alpha * dt,
                      yvt.reshape(-1), Nr, Nvt, Nx)
        return yx.reshape(self.dof_v, Nx)

    def rhs_mvec(self, E, x, dt):
        """bte_1d3v_solver.py:1902-1931: x-shift(x) - dt tau (n0 np0 C x + E A_E x), the
        v-space term zeroed at the inflow points."""
        from .banded import banded_tri_mv
        p = self.params
        Nr, Nvt, Nx = self.Nr, self.Nvt, self.Nx
        x = x.reshape((self.dof_v, Nx))
        idxEp, idxEn = E >= 0, E < 0
        yE = np.zeros_like(x)
        yE[:, idxEp] = self.op_adv_v_Ep @ x[:, idxEp]
        yE[:, idxEn] = self.op_adv_v_En @ x[:, idxEn]
        yv = dt * p.tau * (p.n0 * p.np0 * (self.op_col_en @ x) + E * yE)
        yv = yv.reshape((Nr, Nvt, Nx))
        yv[:, self.idx_np_vt, 0] = 0.0
        yv[:, self.idx_sp_vt, -1] = 0.0
        z = x.reshape((Nr, Nvt, Nx))
        yx = np.zeros_like(z)
        for sel, low in ((self.idx_np_vt, True), (self.idx_sp_vt, False)):
            d = self.bte_x_shift_diag[:, sel].reshape(-1, p.px + 1, Nx)
            yx[:, sel, :] = banded_tri_mv(d, z[:, sel].reshape(-1, Nx), low).reshape(Nr, -1, Nx)
        return (yx.reshape(self.dof_v, Nx) - yv.reshape(self.dof_v, Nx)).reshape(-1)

    def rhs_matrix(self, E, dt):
        """The dense matrix of rhs_mvec, column by column (small sizes only)."""
        n = self.dof_v * self.Nx
        A = np.zeros((n, n))
        e = np.zeros(n)
        for j in range(n):
            e[j] = 1.0
            A[:, j] = self.rhs_mvec(E, e, dt)
            e[j] = 0.0
        return A

    def rhs_vector(self, v):
        """bte_1d3v_solver.py:1897-1900: v with the inflow values zeroed."""
        b = np.copy(v.reshape((self.Nr, self.Nvt, self.Nx)))
        b[:, self.idx_np_vt, 0] = 0.0
        b[:, self.idx_sp_vt, -1] = 0.0
        return b.reshape(-1)

    def step(self, E, v, dt, gmres_pc_vr=1.0, x0=None):
        """bte_1d3v_solver.py:1885-1974: right-preconditioned restarted GMRES on rhs_mvec with
        step_op_split as the preconditioner; raises on non-convergence."""
        p = self.params
        b = self.rhs_vector(v)
        A = lambda x: self.rhs_mvec(E, x, dt)
        M = lambda x: self.step_op_split(E, x.reshape(self.dof_v, self.Nx), dt, gmres_pc_vr).reshape(-1)
        y, info = gmres_pcr(A, b, x0=x0, rtol=p.rtol, atol=p.atol, M=M, restart=p.gmres_rsrt,
                            maxiter=p.gmres_maxiter)
        res = np.linalg.norm(A(y) - b)
        self.gmres_info = {"status": info, "residual_abs": res, "residual_rel": res / np.linalg.norm(b)}
        if info != 0:
            raise RuntimeError("GMRES solver failed! iterations =%d ||res|| = %.4E" % (info, res))
        return y.reshape(self.dof_v, self.Nx)
