"""Fully implicit 1D3V Boltzmann solver on the finite-volume / ordinate velocity grid, on the
MI355X: bte_1d3v (BESolver/python/bte_1d3v_solver.py:218) with vspace_type FVM,
solver_type FULLY_IMPLICIT, pc_type XVTVRC and xspace_type BE_UPW_FD.

Same names, arguments and units as the reference: params from the [bte] table of a TOML
file (:77-216), initial_condition (:1277-1326), init(dt, gmres_pc_cv, gmres_pc_ax)
(:1333-1502), step_op_split(E, v, dv, time, dt, type, verbose, alpha) (:1745-1870),
step(E, v, dv, time, dt, verbose, gmres_pc_vr, gmres_pc_vt, pc_left) (:1872-2041),
compute_qoi (:1179-1216), normalize_edf (:2055-2061).  E is the field at the x points in
V/m; v is the state (Nr*Nvt, Nx), radial point r, ordinate t, point x, x fastest, a float64
device tensor.

Host work is set-up only, once: the grids, the stencil weights (Fornberg's recursion on
the reference's one-sided stencils), the strong-form collision operator in its structured
form (gain matrix G, Nr x Nr with two entries per process per row, and the loss diagonal),
the band tables of the sweeps and the collision solve's Nr x Nr correction matrix.  Each
GMRES iteration runs the step operator (bz_fvm_rhs_mvec: two launches) and the XVTVRC
preconditioner (bz_fvm_precond_xvtvrc: five launches) on the device, inside the device
GMRES of krylov.gmres_pcr.  No dense (Nr Nvt)^2 matrix is formed: the reference inverts
I - dt tau n0 np0 C densely (PmatC) up to dof_v = 65536, and uses a BSR triangular solve
beyond; the structured solve here is exact at every size.
"""
import ctypes
import math

import numpy as np
import scipy.constants as sc
import torch

from . import _lib
from .basis import BSpline
from .cross_section import read_cross_section_data
from .krylov import DeviceGMRES, gmres_pcr

QE = sc.e
ME = sc.electron_mass
KB = sc.Boltzmann
C_GAMMA = math.sqrt(2.0 * QE / ME)
TEMP_K_1EV = sc.electron_volt / KB


class solver_type:
    """bte_1d3v_solver.py:56-60 (this library runs FULLY_IMPLICIT)."""
    XVX_STRANG_SPLIT = 0
    FULLY_IMPLICIT = 1
    SEMILAGRANGIAN_IMPLICIT = 2


class pc_type:
    """bte_1d3v_solver.py:61-68 (this library implements XVTVRC and NONE)."""
    XVTVRC = 0
    XVX_SOLVE = 1
    CXVTVR = 2
    XCVTVR = 3
    NONE = 4


class vspace_discretization:
    FVM = 0
    SPECTRAL_BSPH = 1


class xspace_discretization:
    BE_UPW_FD = 0
    BE_CHEB = 1


def _load_toml(path):
    try:
        import tomllib as tl
    except ImportError:
        import tomli as tl
    with open(path, "rb") as f:
        return tl.load(f)


class params:
    """The [bte] table (bte_1d3v_solver.py:77-216).  Keys the reference reads are read the
    same way; vspace_type must be 0 (FVM), xspace_type 0 (BE_UPW_FD) and solver_type
    "fully_implicit" here.  Without a file, `tp` gives the table as a dict."""

    def __init__(self, par_file=None, tp=None):
        if tp is None:
            tp = _load_toml(par_file)["bte"]
        self.L = 0.5 * tp["L"]
        self.V0 = tp["V0"]
        self.f = tp["freq"]
        self.tau = 1.0 / self.f
        self.qe, self.eps0, self.kB = QE, sc.epsilon_0, KB
        self.ev_to_K = sc.electron_volt / KB
        self.me = ME
        self.Tg = tp["Tg"]
        self.p0 = tp["p0"]
        self.n0 = (self.p0 * sc.torr / (KB * self.Tg)) if self.Tg != 0 else 3.22e22 * self.p0
        self.np0 = 8e16
        self.n0 = self.n0 / self.np0
        self.collisions = tp["collisions"]
        self.sp_order = tp["sp_order"]
        self.spline_qpts = tp["spline_qpts"]
        self.Nr = tp["Nr"]
        self.l_max = tp.get("lmax", 0)
        self.Np = tp["Np"]
        self.Nvt = tp["Nvt"]
        self.dt = tp["dt"]
        self.T = tp.get("cycles", 1)
        self.Te = tp["Te"]
        self.ev_max = tp["ev_max"]
        self.ev_extend = tp.get("ev_extend", 0)
        self.atol = tp["atol"]
        self.rtol = tp["rtol"]
        self.max_iter = tp.get("max_iter", 300)
        self.gmres_rsrt = tp["gmres_rsrt"]
        self.gmres_rsrt_iter = tp.get("gmres_rsrt_iter", 10)
        self.gmres_maxiter = self.gmres_rsrt_iter * self.gmres_rsrt
        self.verbose = tp.get("verbose", 0)
        self.ic_file = tp.get("ic_file", "")
        self.xspace_type = xspace_discretization.BE_UPW_FD if tp.get("xspace_type", 0) == 0 else xspace_discretization.BE_CHEB
        self.vspace_type = vspace_discretization.FVM if tp.get("vspace_type", 0) == 0 else vspace_discretization.SPECTRAL_BSPH
        self.pvr = tp.get("pvr", 3)
        self.pvt = tp.get("pvt", 3)
        self.px = tp.get("px", 1)
        self.adaptive_grid = bool(tp.get("adaptive_grid", 1))
        self.pc_type = {"none": pc_type.NONE}.get(tp.get("pc_type", "xvtvrc"), pc_type.XVTVRC)
        if tp.get("pc_type", "xvtvrc") not in ("xvtvrc", "none"):
            raise NotImplementedError("pc_type %s: XVTVRC and none are implemented" % tp["pc_type"])
        st = tp.get("solver_type", "fully_implicit")
        if st != "fully_implicit":
            raise NotImplementedError("solver_type %s: this solver implements fully_implicit" % st)
        self.solver_type = solver_type.FULLY_IMPLICIT
        if self.vspace_type != vspace_discretization.FVM or self.xspace_type != xspace_discretization.BE_UPW_FD:
            raise NotImplementedError("the fully implicit solver runs on vspace_type 0 (FVM), xspace_type 0")


# ----------------------------------------------------------------------------- stencils

def fornberg_weights(z, x, m):
    """Weights c[j, k] with d^k f/dx^k (z) ~ sum_j c[j, k] f(x_j), k <= m, exact for
    polynomials of degree < len(x) (B. Fornberg, Math. Comp. 51 (1988) 699-706)."""
    n = len(x)
    c = np.zeros((n, m + 1))
    c1, c4 = 1.0, x[0] - z
    c[0, 0] = 1.0
    for i in range(1, n):
        mn = min(i, m)
        c2, c5, c4 = 1.0, c4, x[i] - z
        for j in range(i):
            c3 = x[i] - x[j]
            c2 *= c3
            if j == i - 1:
                for k in range(mn, 0, -1):
                    c[i, k] = c1 * (k * c[i - 1, k - 1] - c5 * c[i - 1, k]) / c2
                c[i, 0] = -c1 * c5 * c[i - 1, 0] / c2
            for k in range(mn, 0, -1):
                c[j, k] = (c4 * c[j, k] - k * c[j, k - 1]) / c3
            c[j, 0] = c4 * c[j, 0] / c3
        c1 = c2
    return c


def _d1(x, at):
    """First-derivative weights on the points x at x[at]."""
    return fornberg_weights(x[at], x, 1)[:, 1]


def upwinded_dx(x, sw, dir):
    """The reference's one-sided first derivative (mesh.py:20-56): sw-point stencils ending
    at row i ("L") or starting at it ("R"), shorter next to the start / end."""
    n = len(x)
    D = np.zeros((n, n))
    if dir == "L":
        for i in range(1, n):
            lo = max(0, i - (sw - 1))
            D[i, lo:i + 1] = _d1(x[lo:i + 1], i - lo)
    else:
        for i in range(0, n - 1):
            hi = min(n, i + sw)
            D[i, i:hi] = _d1(x[i:hi], 0)
    return D


def upwinded_dvt(x, sw, dir):
    """mesh.py:80-106 (use_cdx_internal=False): each half of the ordinates on its own, the
    halves joined by a two-point row."""
    n = len(x)
    mp = n // 2
    D = np.zeros((n, n))
    D[:mp, :mp] = upwinded_dx(x[:mp], sw, dir)
    D[mp:, mp:] = upwinded_dx(x[mp:], sw, dir)
    if dir == "L":
        D[mp] = 0.0
        D[mp, mp - 1:mp + 1] = _d1(x[mp - 1:mp + 1], 1)
    else:
        D[mp - 1] = 0.0
        D[mp - 1, mp - 1:mp + 1] = _d1(x[mp - 1:mp + 1], 0)
    return D


def row_bands(D, nd, lower):
    """[d, i] = D[i, i -/+ d] (zero outside): the row-wise bands the step operator reads."""
    n = D.shape[0]
    B = np.zeros((nd, n))
    for d in range(nd):
        i = np.arange(d, n) if lower else np.arange(0, n - d)
        B[d, i] = D[i, i - d] if lower else D[i, i + d]
    return B


def column_bands(D, nd, lower):
    """[k, j] = D[j + k, j] (lower) or D[j - k, j] (upper): the sweeps' layout
    (bte_1d3v_solver.py:1440-1450)."""
    n = D.shape[0]
    B = np.zeros((nd, n))
    for k in range(nd):
        j = np.arange(0, n - k) if lower else np.arange(k, n)
        B[k, j] = D[j + k, j] if lower else D[j - k, j]
    return B


def _xs(rec, ev):
    """Linear interpolation of a cross-section table, constant beyond its ends
    (collisions.py:79)."""
    return np.interp(ev, rec["energy"], rec["cross section"], left=rec["cross section"][0],
                     right=rec["cross section"][-1])


# ----------------------------------------------------------------------------- the solver

class gmres_counter:
    """Iteration count and relative residual history of one GMRES solve (the reference's
    utils.gmres_counter with store_residuals=True)."""

    def __init__(self):
        self.niter = 0
        self.residuals = []

    def __call__(self, rk):
        self.niter += 1
        self.residuals.append(float(rk))


class bte_1d3v:
    def __init__(self, args):
        self.params = p = args.params if getattr(args, "params", None) is not None else params(args.par_file)
        self.dev = _lib.require_device()
        self.qe, self.me, self.kB, self.c_gamma = QE, ME, KB, C_GAMMA
        self.xp_module = torch
        Nvt = p.Nvt
        if Nvt % 2:
            raise ValueError("Nvt must be even")
        if p.adaptive_grid:
            self.xp = -np.cos(np.pi * np.linspace(0, p.Np - 1, p.Np) / (p.Np - 1))
        else:
            self.xp = np.linspace(-1, 1, p.Np)
        self.bs_Te = p.Te
        self.bs_vth = vth = math.sqrt(2.0 * KB * (p.Te * QE / KB) / ME)
        self.ev_lim = (0, p.ev_max)
        self.cross_section_data = read_cross_section_data(p.collisions)
        self.coll_list = list(self.cross_section_data.values())
        self.num_collisions = len(self.coll_list)
        mbym = [r["mass_ratio"] for r in self.coll_list if r["type"] in ("ELASTIC", "EFFECTIVE")]
        self._mByM = mbym[0] if mbym else None
        k_domain = (math.sqrt(self.ev_lim[0]) * C_GAMMA / vth, math.sqrt(self.ev_lim[1]) * C_GAMMA / vth)
        bb = BSpline(k_domain, p.sp_order, p.Nr + 1, sig_pts=None, knots_vec=None,
                     extend_domain=(p.ev_extend == 1), extend_domain_with_log=(p.ev_extend == 2))
        self._basis_p = bb
        self.xp_vr, self.xp_vr_qw = bb.Gauss_Pn(p.spline_qpts * bb._num_knot_intervals)
        gx, gw = np.polynomial.legendre.leggauss(Nvt // 2)
        self.xp_vt = np.arccos(np.append(0.5 * gx - 0.5, 0.5 * gx + 0.5))
        self.xp_vt_qw = np.append(0.5 * gw, 0.5 * gw)
        assert self.xp_vr[0] > bb._domain[0] and self.xp_vr[-1] < bb._domain[1]
        self.dof_vr, self.dof_vt, self.dof_x = len(self.xp_vr), Nvt, p.Np
        self.dof_v = self.dof_vr * Nvt
        self._assemble()
        self.is_init = False
        # GMRES with the Arnoldi step on the device (krylov.DeviceGMRES, graph-replayed); False
        # runs the reference's loop (krylov.gmres_pcr, true residual after every step)
        self.use_device_gmres = True
        # Gram-Schmidt passes of the device GMRES: one (classical) by default, the cost of the
        # reference's one modified pass.  The steps per solve equal those of two passes on the
        # measured 671 x 16 x 100 runs, and a cycle's solution is taken only after its true
        # residual passes the tolerance
        self.gs_passes = 1
        self._dg, self._dg_key = None, None
        self._Ebuf = torch.zeros(self.dof_x, dtype=torch.float64, device=self.dev)

    # -------------------------------------------------------------- host set-up
    def _collision(self):
        """The strong-form operator of collision_operator_spherical.py:1334-1419, summed over
        the processes, as (G, loss): (C f)[r, t] = (G s)[r] - loss[r] f[r, t] with
        s[r'] = sum_t 2 pi w_t f[r', t]."""
        vr, vth = self.xp_vr, self.bs_vth
        Nr = len(vr)
        rmax = float(np.sum(self.xp_vr_qw))
        G = np.zeros((Nr, Nr))
        loss = np.zeros(Nr)
        for rec in self.coll_list:
            th = rec["threshold"]
            typ = rec["type"]
            if typ == "ELASTIC":
                v_scale = math.sqrt(1 - 2 * self._mByM)
                vpre = vr / v_scale
                detJ = np.full(Nr, (1 / v_scale) ** 3)
            elif typ == "EXCITATION":
                vpre = np.sqrt(vr ** 2 + (C_GAMMA / vth) ** 2 * th)
                detJ = vpre / vr
            elif typ == "IONIZATION":
                vpre = np.sqrt(2 * vr ** 2 + (C_GAMMA / vth) ** 2 * th)
                detJ = 4.0 * vpre / vr
            else:
                raise NotImplementedError("collision type %s on the FVM grid" % typ)
            dcs = (1 / (4 * np.pi)) * vr * vth * _xs(rec, (vr * vth / C_GAMMA) ** 2)
            dcs_pre = (1 / (4 * np.pi)) * vpre * vth * _xs(rec, (vpre * vth / C_GAMMA) ** 2)
            # hat-function interpolation at the pre-collision speed; linear to zero at the end
            # of the domain beyond the last point
            k = np.clip(np.searchsorted(vr, vpre, side="left"), 1, Nr - 1)
            inside = (vpre >= vr[0]) & (vpre <= vr[-1])
            wgt = (vpre - vr[k - 1]) / (vr[k] - vr[k - 1])
            rows = np.nonzero(inside)[0]
            G[rows, k[rows] - 1] += (dcs_pre * detJ)[rows] * (1.0 - wgt[rows])
            G[rows, k[rows]] += (dcs_pre * detJ)[rows] * wgt[rows]
            tail = np.nonzero((vpre > vr[-1]) & (vpre <= rmax))[0]
            G[tail, Nr - 1] += (dcs_pre * detJ)[tail] * (rmax - vpre[tail]) / (rmax - vr[-1])
            loss += 4 * np.pi * dcs
        return G, loss

    def _assemble(self):
        p = self.params
        vr, vt, vth = self.xp_vr, self.xp_vt, self.bs_vth
        Nr, Nvt, Nx = self.dof_vr, self.dof_vt, self.dof_x
        self.G, self.loss = self._collision()
        ev_fac = 0.5 * ME * (2 / 3 / KB) / TEMP_K_1EV
        self.op_mass = 2 * np.pi * np.kron(vr ** 2 * self.xp_vr_qw, self.xp_vt_qw)
        self.op_temp = 2 * np.pi * vth ** 2 * np.kron(vr ** 4 * self.xp_vr_qw, self.xp_vt_qw) * ev_fac
        self.op_rate = [(2 * np.pi) * (2 * vth ** 4 / C_GAMMA ** 3) *
                        np.kron(vr ** 3 * self.xp_vr_qw * _xs(rec, (vr * vth / C_GAMMA) ** 2), self.xp_vt_qw)
                        for rec in self.coll_list]
        self.op_adv_x_d = vr * vth * (p.tau / p.L)
        # radial and angular derivatives of the step operator (spec_spherical.py:905-923)
        dom = self._basis_p._domain
        DvrL, DvrR = upwinded_dx(vr, p.pvr + 1, "L"), upwinded_dx(vr, p.pvr + 1, "R")
        DvrL[0, :] = 0.0
        DvrR[-1, :] = 0.0
        DvrR[-1, -1] = -1 / (dom[1] - vr[-1])
        DvtL, DvtR = upwinded_dvt(vt, p.pvt + 1, "L"), upwinded_dvt(vt, p.pvt + 1, "R")
        DvtL[0, :] = 0.0
        DvtR[-1, :] = 0.0
        self._DvtL, self._DvtR = DvtL, DvtR
        self._t = {}
        d = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(self.dev, dt)
        self._t.update(cos_vt=d(np.cos(vt)), sin_vt=d(np.sin(vt)), vr=d(vr), wq=d(2 * np.pi * self.xp_vt_qw),
                       loss=d(self.loss), adx=d(self.op_adv_x_d),
                       dvrL=d(row_bands(DvrL, p.pvr + 1, True)), dvrR=d(row_bands(DvrR, p.pvr + 1, False)),
                       dvtL=d(row_bands(DvtL, p.pvt + 1, True)), dvtR=d(row_bands(DvtR, p.pvt + 1, False)),
                       axvr=d(vth * (p.tau / p.L) * vr))
        gp, gc, gv = [0], [], []
        for r in range(Nr):
            nz = np.nonzero(self.G[r])[0]
            gc.extend(nz.tolist())
            gv.extend(self.G[r, nz].tolist())
            gp.append(len(gc))
        self._t.update(g_ptr=d(np.array(gp), torch.int32), g_col=d(np.array(gc if gc else [0]), torch.int32),
                       g_val=d(np.array(gv if gv else [0.0])))
        self._work = torch.empty(2 * Nr * Nx, dtype=torch.float64, device=self.dev)
        # the QoI weight vectors [mass, temp, rate_0, ...] in groups of three (zero padded) for
        # the one-pass weighted sums of bz_glow_moments
        wts = [self.op_mass, self.op_temp] + list(self.op_rate)
        wts += [np.zeros(self.dof_v)] * (-len(wts) % 3)
        self._qoi_w = [d(np.stack(wts[i:i + 3])) for i in range(0, len(wts), 3)]
        self._qoi_split = 64

    # -------------------------------------------------------------- state and QoIs
    def initial_condition(self, type=0):
        """bte_1d3v_solver.py:1277-1326 (FVM, type 0, no IC file): ne(x) times a Maxwellian at
        Te, the same at every ordinate."""
        p = self.params
        if type != 0 or p.ic_file:
            raise NotImplementedError("initial condition type 0 without an IC file")
        xx = p.L * (self.xp + 1)
        ne = 1e6 * (1e7 + 1e9 * (1 - 0.5 * xx / p.L) ** 2 * (0.5 * xx / p.L) ** 2) / p.np0
        v_ratio = C_GAMMA * math.sqrt(p.Te) / self.bs_vth
        hv = (1 / np.sqrt(np.pi) ** 3) * np.exp(-((self.xp_vr / v_ratio) ** 2)) / v_ratio ** 3
        v = np.repeat(hv, self.dof_vt)[:, None] * ne[None, :]
        return torch.from_numpy(np.ascontiguousarray(v)).to(self.dev)

    def _weighted_sums(self, v):
        """(2 + processes, Nx) sums over the state rows of [mass, temp, rate_k] x v: one pass of
        bz_glow_moments per three weight vectors (64 row splits), then the split sums."""
        v = v.reshape(self.dof_v, self.dof_x).contiguous()
        out = []
        for W in self._qoi_w:
            part = torch.empty((self._qoi_split, 3, self.dof_x), dtype=torch.float64, device=self.dev)
            _lib.call("bz_glow_moments", v, self.dof_v, self.dof_x, W, self._qoi_split, part)
            out.append(part.sum(0))
        return torch.cat(out)[:2 + self.num_collisions]

    def normalize_edf(self, v):
        m0 = self._weighted_sums(v)[0]
        return v / (m0 * (2 * (self.bs_vth / C_GAMMA) ** 3))

    def compute_qoi(self, v, time=0.0, dt=0.0):
        """bte_1d3v_solver.py:1179-1216 (FVM): ne, Te and the rate coefficients per x point.
        The rates of the normalised EDF are the rate sums of v over m0 2 (vth / c_gamma)^3."""
        S = self._weighted_sums(v)
        m0 = S[0]
        Te = S[1] / m0
        rates = (S[2:] / (m0 * (2 * (self.bs_vth / C_GAMMA) ** 3))).T
        return {"rates": rates, "mu": 0, "D": 0, "ne": m0 * self.params.np0, "Te": Te}

    # -------------------------------------------------------------- init
    def init(self, dt, gmres_pc_cv=1.0, gmres_pc_ax=1.0):
        """bte_1d3v_solver.py:1333-1502 for FVM: the collision solve with step
        gmres_pc_cv dt, the x shift of the step operator with step gmres_pc_ax dt, the band
        tables of the sweeps."""
        p = self.params
        vr, vt = self.xp_vr, self.xp_vt
        Nr, Nvt, Nx = self.dof_vr, self.dof_vt, self.dof_x
        d = lambda a, dt_=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to(self.dev, dt_)
        # collision solve: (I - c C)^-1 b = b / A + H s_b, A = 1 + c loss,
        # H = diag(c / A) G (diag(A) - c W G)^-1, W = sum_t 2 pi w_t
        c = gmres_pc_cv * dt * p.tau * p.n0 * p.np0
        A = 1.0 + c * self.loss
        W = float(np.sum(2 * np.pi * self.xp_vt_qw))
        M = np.linalg.inv(np.diag(A) - c * W * self.G)
        H = (c / A)[:, None] * (self.G @ M)
        nzc = np.abs(H) > 0
        h_lo = np.where(nzc.any(axis=1), nzc.argmax(axis=1), Nr).astype(np.int32)
        # x derivatives: the step operator's (BE_UPW_FD, :801-817) and the sweep's (:1416-1425)
        DxL, DxR = upwinded_dx(self.xp, p.px + 1, "L"), upwinded_dx(self.xp, p.px + 1, "R")
        DxL[0, :] = 0.0
        DxR[-1, :] = 0.0
        # sweep bands (:1412-1450): the radial outflow row uses the unextended domain end
        kd = self._basis_p._kdomain
        DvrL, DvrR = upwinded_dx(vr, p.pvr + 1, "L"), upwinded_dx(vr, p.pvr + 1, "R")
        DvrL[0, :] = 0.0
        DvrR[-1, :] = 0.0
        DvrR[-1, -1] = -1 / (kd[1] - vr[-1])
        sin = np.diag(np.sin(vt))
        self._t.update(ainv=d(1.0 / A), H=d(H), h_lo=d(h_lo, torch.int32),
                       dxL=d(row_bands(DxL, p.px + 1, True)), dxR=d(row_bands(DxR, p.px + 1, False)),
                       vr_l2r=d(column_bands(DvrL, p.pvr + 1, True)), vr_r2l=d(column_bands(DvrR, p.pvr + 1, False)),
                       vt_l2r=d(column_bands(sin @ self._DvtL, p.pvt + 1, True)),
                       vt_r2l=d(column_bands(sin @ self._DvtR, p.pvt + 1, False)),
                       x_l2r=d(column_bands(DxL, p.px + 1, True)), x_r2l=d(column_bands(DxR, p.px + 1, False)))
        t = self._t
        P = lambda k: ctypes.c_void_p(t[k].data_ptr())
        self._op = _lib.Fvm1d3v(
            Nr, Nvt, Nx, p.pvr + 1, p.pvt + 1, p.px + 1,
            P("cos_vt"), P("sin_vt"), P("vr"), P("wq"), P("loss"), P("g_ptr"), P("g_col"), P("g_val"),
            P("dvrL"), P("dvrR"), P("dvtL"), P("dvtR"), P("dxL"), P("dxR"), P("adx"),
            gmres_pc_ax * dt, p.tau * p.n0 * p.np0, p.tau, (1 / self.bs_vth) * (QE / ME),
            (p.tau * p.qe / p.me / self.bs_vth),
            P("vr_l2r"), P("vr_r2l"), P("vt_l2r"), P("vt_r2l"), P("x_l2r"), P("x_r2l"), P("axvr"), P("ainv"), P("H"),
            P("h_lo"), ctypes.c_void_p(self._work.data_ptr()))
        self.adv_setup_dt = gmres_pc_ax * dt
        self.is_init = True
        self._dg, self._dg_key = None, None        # captured launches hold the old tables

    def _device_gmres(self, dt, alpha):
        key = (float(dt), float(alpha))
        if self._dg is None or self._dg_key != key:
            op, E = self._op, self._Ebuf
            mv = lambda src, dst: _lib.call("bz_fvm_rhs_mvec", op, E, key[0], src, dst)
            pc = lambda src, dst: _lib.call("bz_fvm_precond_xvtvrc", op, E, key[1] * key[0], src, dst)
            self._dg = DeviceGMRES(self.dof_v * self.dof_x, self.params.gmres_rsrt, self.dev, mv, pc,
                                   gs_passes=self.gs_passes)
            self._dg_key = key
        return self._dg

    # -------------------------------------------------------------- step
    def _field(self, E):
        E = torch.as_tensor(E, dtype=torch.float64, device=self.dev).reshape(-1).contiguous()
        if E.numel() != self.dof_x:
            raise ValueError("E must hold one value per x point")
        return E

    def rhs_mvec(self, E, x, dt):
        """The step operator (bte_1d3v_solver.py:1902-1931) on a flat device vector."""
        y = torch.empty_like(x)
        _lib.call("bz_fvm_rhs_mvec", self._op, E, dt, x.contiguous(), y)
        return y

    def step_op_split(self, E, v, dv, time, dt, type=pc_type.XVTVRC, verbose=0, alpha=1.0):
        """XVTVRC (bte_1d3v_solver.py:1745-1870): collisions, then vr, vt and x sweeps."""
        if type != pc_type.XVTVRC:
            raise NotImplementedError("preconditioner type not implemented")
        if not self.is_init:
            raise RuntimeError("init(dt) must be called before step_op_split")
        E = self._field(E)
        b = v.reshape(-1).contiguous()
        y = torch.empty_like(b)
        _lib.call("bz_fvm_precond_xvtvrc", self._op, E, alpha * dt, b, y)
        return y.reshape(self.dof_v, self.dof_x)

    def step(self, E, v, dv, time, dt, verbose=0, gmres_pc_vr=1.0, gmres_pc_vt=1.0, pc_left=False):
        """One backward-Euler step of the full 1D3V operator (bte_1d3v_solver.py:1885-1974):
        right-preconditioned restarted GMRES from x0 = 0 on the step operator with the
        inflow values of v zeroed; raises RuntimeError if GMRES does not converge."""
        p = self.params
        if pc_left:
            raise NotImplementedError("left preconditioning (gmres_pcl) is not implemented")
        if not self.is_init:
            raise RuntimeError("init(dt) must be called before step")
        E = self._field(E)
        Nr, Nvt, Nx = self.dof_vr, self.dof_vt, self.dof_x
        mp = Nvt // 2
        brhs = v.reshape(Nr, Nvt, Nx).clone()
        brhs[:, mp:, 0] = 0.0            # theta <= pi/2 enters at x = -1
        brhs[:, :mp, -1] = 0.0           # theta > pi/2 enters at x = 1
        brhs = brhs.reshape(-1)
        A = lambda x: self.rhs_mvec(E, x, dt)
        if p.pc_type == pc_type.XVTVRC:
            M = lambda x: self.step_op_split(E, x, None, time, dt, pc_type.XVTVRC, 0, gmres_pc_vr).reshape(-1)
        else:
            M = None
        cb = gmres_counter()
        if self.use_device_gmres and p.pc_type == pc_type.XVTVRC:
            dg = self._device_gmres(dt, gmres_pc_vr)
            self._Ebuf.copy_(E)
            cycles = max(1, -(-p.gmres_maxiter // p.gmres_rsrt))
            try:
                y, iters = dg.solve(brhs, p.atol, p.rtol, cycles)
                status = 0
            except RuntimeError:
                y, status = dg.t.clone(), cycles * p.gmres_rsrt
            cb.niter = iters if status == 0 else status
        else:
            y, status = gmres_pcr(A, brhs, x0=None, rtol=p.rtol, atol=p.atol, M=M, restart=p.gmres_rsrt,
                                  maxiter=p.gmres_maxiter, callback=cb)
        res = torch.linalg.vector_norm(A(y) - brhs)
        nb = torch.linalg.vector_norm(brhs)
        self.gmres_info = {"status": status, "residual_abs": float(res), "residual_rel": float(res / nb),
                           "gmres_callback": cb}
        if status != 0:
            raise RuntimeError("%08d GMRES solver failed! iterations =%d  ||res|| = %.4E ||res||/||b|| = %.4E"
                               % (int(time / dt), status, float(res), float(res / nb)))
        if verbose == 1:
            print("%08d ||res|| = %.4E ||res||/||b|| = %.4E" % (int(time / dt), float(res), float(res / nb)))
        return y.reshape(self.dof_v, Nx)
