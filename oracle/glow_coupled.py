"""Oracle: the plasma coupling of the 1D3V glow-discharge step (FP64, numpy).
TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates BESolver/python:
  * collocation_op.cheb_collocation_1d (collocation_op.py:8-51): Chebyshev-Gauss-Lobatto
    points xp_k = -cos(pi k / (Np-1)), the first/second derivative matrices from the
    Chebyshev Vandermonde matrix, LpD (identity boundary rows, Lp interior rows) and its
    inverse;
  * glow1d_utils.parameters (glow1d_utils.py:26-130): the default branch (L = 1.27 cm half
    gap, 100 V, 13.56 MHz, 1 Torr, 300 K) or a [glow_1d] table, and the nondimensional
    scaling of n0, D and mu by tau / L^2 and V0 tau / L^2;
  * glow1d_boltzmann (glowdischarge_boltzmann_1d.py): solve_poisson (:1005-1019),
    bte_eedf_normalization (:1021-1030), bte_to_fluid (:1050-1076), fluid_to_bte
    (:1078-1086), rhs_fluid (:1088-1149) and rhs_fluid_jacobian (:1151-1218) with the weak
    ion boundary flux and the frozen field (heavies_freeze_E), the backward-Euler step_fluid
    (:1318-1372) through glow1d_utils.newton_solver (glow1d_utils.py:131-207), step_bte
    (:2030-2071), the Strang step (:2214-2241) and initialize's default state (:541-655).
    The electron v-step is solved exactly per point (oracle/krylov.vstep_exact), the limit of
    the reference's GMRES.

Pinned by tests/test_oracle_glow_coupled.py: the derivative matrices are exact on
polynomials of degree < Np and agree with the closed-form Chebyshev differentiation matrix;
the Poisson solve reproduces a manufactured polynomial potential with its boundary values;
the Jacobian equals the directional difference of the (affine) right-hand side; the fluid
step satisfies the backward-Euler equation and conserves ions up to the volume source and
the boundary fluxes (Clenshaw-Curtis weights integrate D_p exactly); the moments of a
Maxwellian give its density and temperature.
"""
import math

import numpy as np
import scipy.constants as sc

from . import glow1d as gl
from . import krylov as kr
from . import operators as op


# ----------------------------------------------------------------------------- collocation

def cheb_collocation_1d(Np):
    """(xp, Dp, Lp, LpD, LpD_inv), collocation_op.py:8-51."""
    deg = Np - 1
    xp = -np.cos(np.pi * np.linspace(0, deg, Np) / deg)
    I = np.identity(Np)
    V0p = np.polynomial.chebyshev.chebvander(xp, deg)
    V0pinv = np.linalg.solve(V0p, I)
    V1p = np.array([np.polynomial.Chebyshev.basis(i).deriv(m=1)(xp) for i in range(Np)]).reshape((-1, Np)).T
    V2p = np.array([np.polynomial.Chebyshev.basis(i).deriv(m=2)(xp) for i in range(Np)]).reshape((-1, Np)).T
    Dp = V1p @ V0pinv
    Lp = V2p @ V0pinv
    LpD = np.identity(Np)
    LpD[1:-1, :] = Lp[1:-1, :]
    return xp, Dp, Lp, LpD, np.linalg.solve(LpD, I)


def cheb_diff_closed_form(Np):
    """Textbook Chebyshev differentiation matrix on xp_k = -cos(pi k / (Np-1)) (ascending):
    D_ij = (c_i / c_j) (-1)^(i+j) / (x_i - x_j) off the diagonal, c_0 = c_N = 2, else 1; the
    diagonal from the zero row sums.  Used only to pin cheb_collocation_1d."""
    N = Np - 1
    x = -np.cos(np.pi * np.arange(Np) / N)
    c = np.ones(Np)
    c[0] = c[-1] = 2.0
    c = c * (-1.0) ** np.arange(Np)
    X = x[:, None] - x[None, :]
    D = np.outer(c, 1.0 / c) / (X + np.eye(Np))
    D = D - np.diag(D.sum(axis=1))
    return x, D


# ----------------------------------------------------------------------------- parameters

class Parameters:
    """glow1d_utils.parameters (glow1d_utils.py:26-130).  tp: the [glow_1d] table of a
    parameter file (L is the full gap there) or None for the default branch."""

    def __init__(self, tp=None):
        if tp is not None:
            self.L = 0.5 * tp["L"]
            self.V0 = tp["V0"]
            self.f = tp["freq"]
            self.Tg = tp["Tg"]
            self.p0 = tp["p0"]
            self.Teb = tp.get("Teb")
            self.gamma = tp.get("gamma")
            De, mu_e, Di, mu_i = tp["De"], tp["mu_e"], tp["Di"], tp["mu_i"]
        else:
            self.L = 0.5 * 2.54e-2
            self.V0 = 1e2
            self.f = 13.56e6
            self.Tg = 300.0
            self.p0 = 1.0
            self.Teb = None
            self.gamma = None
            De, mu_e, Di, mu_i = 3.86e22 * 1e2, 9.66e21 * 1e2, 2.07e18 * 1e2, 4.65e19 * 1e2
        self.tau = 1.0 / self.f
        self.n0 = self.p0 * sc.torr / (sc.Boltzmann * self.Tg) if self.Tg != 0 else 3.22e22
        self.np0 = 8e16
        self._De = De / self.n0
        self._mu_e = mu_e / self.n0
        self.Di = Di / self.n0
        self.mu_i = mu_i / self.n0
        self.n0 /= self.np0
        self._De *= self.tau / self.L ** 2
        self.Di *= self.tau / self.L ** 2
        self._mu_e *= self.V0 * self.tau / self.L ** 2
        self.mu_i *= self.V0 * self.tau / self.L ** 2
        self.alpha = self.np0 * self.L ** 2 * sc.e / (sc.epsilon_0 * self.V0)


# ----------------------------------------------------------------------------- fluid pieces

def solve_poisson(LpD_inv, alpha, ne, ni, time):
    """:1005-1019: phi with phi'' = -alpha (ni - ne), phi(-1) = sin(2 pi t), phi(1) = 0."""
    r = -alpha * (ni - ne)
    r[0] = np.sin(2 * np.pi * time)
    r[-1] = 0.0
    return LpD_inv @ r


def rhs_fluid(Dp, mu_i, Di, n0, ki, E, ne, ni):
    """:1088-1149 for the ion density (weak boundary flux, frozen field E)."""
    flux = mu_i * ni * E - Di * (Dp @ ni)
    flux[0] = mu_i[0] * ni[0] * E[0]
    flux[-1] = mu_i[-1] * ni[-1] * E[-1]
    return ki * n0 * ne - Dp @ flux


def rhs_fluid_jacobian(Dp, mu_i, Di, E):
    """:1151-1218: d rhs / d ni with E frozen (E_ni = 0).  The reference broadcasts the
    point arrays over the columns (mu_i * M scales column j by mu_i[j])."""
    Np = Dp.shape[0]
    Imat = np.eye(Np)
    Ji = mu_i * (E * Imat) - Di * Dp
    Ji[0, :] = mu_i[0] * (E[0] * Imat[0, :])
    Ji[-1, :] = mu_i[-1] * (E[-1] * Imat[-1, :])
    return -(Dp @ Ji)


def newton_solver(x, residual, jacobian, atol, rtol, iter_max):
    """glow1d_utils.py:131-207: Newton with a frozen inverse Jacobian and a damping factor
    that restarts from x0 at a quarter of the step while the iteration fails."""
    x0 = x.copy()
    jac_inv = np.linalg.inv(jacobian(x0))
    alpha = 1.0
    r0 = residual(x0)
    norm_rr = norm_r0 = np.linalg.norm(r0)
    with np.errstate(invalid="ignore", divide="ignore"):
        converged = bool((norm_rr / norm_r0 < rtol) or (norm_rr < atol))
    count = 0
    while alpha > 1e-6:
        count = 0
        x = x0
        rr = r0
        while (not converged) and count < iter_max:
            x = x - alpha * (jac_inv @ rr).reshape(x.shape)
            rr = residual(x)
            norm_rr = np.linalg.norm(rr)
            count += 1
            converged = bool((norm_rr / norm_r0 < rtol) or (norm_rr < atol))
            if np.isnan(norm_rr) or norm_rr > norm_r0 * 1e2:
                break
        if not converged:
            alpha *= 0.25
        else:
            break
    return dict(x=x, status=converged, atol=norm_rr, rtol=norm_rr / norm_r0 if norm_r0 else 0.0,
                alpha=alpha, iter=count)


def step_fluid_be(Dp, mu_i, Di, n0, ki, E, u, dt, atol, rtol, iter_max, ion=1, ele=0):
    """:1318-1372 (ts_type BE): u with the ion column advanced by one backward-Euler step."""
    ne = u[:, ele]

    def residual(du):
        return du - dt * rhs_fluid(Dp, mu_i, Di, n0, ki, E, ne, u[:, ion] + du)

    def jacobian(du):
        return np.eye(Dp.shape[0]) - dt * rhs_fluid_jacobian(Dp, mu_i, Di, E)

    info = newton_solver(np.zeros(u.shape[0]), residual, jacobian, atol, rtol, iter_max)
    if not info["status"]:
        raise RuntimeError("fluid step failed: " + repr({k: info[k] for k in ("atol", "rtol", "iter")}))
    u1 = u.copy()
    u1[:, ion] += info["x"]
    return u1, info


# ----------------------------------------------------------------------------- coupled step

class GlowCoupled:
    """The Strang-split glow step of the reference with the oracle's velocity-space pieces.

    ops, g, procs: oracle operators of the radial/angular basis (solver0d.BatchedOps);
    Np Chebyshev points, Nvt ordinates, dt in RF periods, param: Parameters.
    ion_proc: index of the ionisation process in procs (the reference's op_rate[1] for its
    two-process cross-section files)."""

    def __init__(self, ops, g, procs, Np, Nvt, dt, param, ion_proc, atol=1e-10, rtol=1e-10, max_iter=1000):
        self.ops, self.g, self.param = ops, g, param
        self.Np, self.Nvt, self.dt = Np, Nvt, dt
        self.atol, self.rtol, self.max_iter = atol, rtol, max_iter
        self.xp, self.Dp, self.Lp, self.LpD, self.LpD_inv = cheb_collocation_1d(Np)
        self.mu_i = np.full(Np, param.mu_i)
        self.Di = np.full(Np, param.Di)
        theta, w = gl.ordinates(Nvt)
        self.po2sh, self.psh2o = gl.sph_ops(g.lmax, theta, w)
        self.xstep = gl.XStep(gl.adv_x_operator(g, g.vth * param.tau / param.L), self.xp, theta, 0.5 * dt)
        self.K0 = param.n0 * param.np0 * (ops.c_en.sum(axis=0) + param.Tg * ops.c_gT)
        self.A = ops.adv
        ns = g.num_sh
        self.mass_l0 = ops.mass[0::ns]
        self.temp_l0 = (op.temp_op(g) * g.vth ** 2 * op.TEMP_EV_FAC)[0::ns]
        self.rate_ion = op.rate_op(g, procs[ion_proc])
        self.mm_fac = math.sqrt(4.0 * math.pi)
        self.c_gamma = op.C_GAMMA
        self.ki = np.zeros(Np)
        self.bs_E = np.zeros(Np)

    def ords_to_sph(self, v):
        return gl.ords_to_sph(self.po2sh, v, self.g.num_p, self.Nvt, self.Np)

    def sph_to_ords(self, v):
        return gl.sph_to_ords(self.psh2o, v, self.g.num_p, self.g.num_sh, self.Np)

    def initialize(self, Te):
        """:541-655, default initial state: ne = ni = 1e6 (1e7 + 1e9 s^2 (1-s)^2) / np0 with
        s = (xp+1)/2, Te everywhere, a unit-mass Maxwellian at Te on l = 0 per point."""
        p, g = self.param, self.g
        xx = p.L * (self.xp + 1)
        u = np.zeros((self.Np, 3))
        prof = 1e6 * (1e7 + 1e9 * (1 - 0.5 * xx / p.L) ** 2 * (0.5 * xx / p.L) ** 2) / p.np0
        u[:, 0] = prof
        u[:, 1] = prof
        u[:, 2] = Te
        h = op.maxwellian_init(g, self.ops.Minv)
        h = h / (self.ops.mass @ h)
        v_col = self.sph_to_ords(np.repeat(h.reshape(-1, 1), self.Np, axis=1))
        return u, v_col * u[:, 0][None, :]

    def bte_to_fluid(self, u, v):
        """:1050-1076: ne, Te and the ionisation rate ki (per RF period, per np0) from v."""
        ns = self.g.num_sh
        v_lm = self.ords_to_sph(v)
        v0 = v_lm[0::ns, :]
        u = u.copy()
        u[:, 0] = self.mass_l0 @ v0
        u[:, 2] = (self.temp_l0 @ v0) / u[:, 0]
        scale = (self.mass_l0 / self.mm_fac) @ v0 * (2 * (self.g.vth / self.c_gamma) ** 3)
        self.ki = (self.rate_ion @ (v0 / scale)) * self.param.np0 * self.param.tau
        return u

    def field(self, u, time):
        """Dp (-phi): the nondimensional field of the fluid step (:2221-2222)."""
        return self.Dp @ (-solve_poisson(self.LpD_inv, self.param.alpha, u[:, 0], u[:, 1], time))

    def fluid_to_bte(self, u, time):
        """:1078-1086: the field of the electron v-step in V/m."""
        phi = solve_poisson(self.LpD_inv, self.param.alpha, u[:, 0], u[:, 1], time)
        self.bs_E = -(self.Dp @ phi) * self.param.V0 / self.param.L
        return self.bs_E

    def step_fluid(self, u, E, dt):
        u1, info = step_fluid_be(self.Dp, self.mu_i, self.Di, self.param.n0, self.ki, E, u, dt, self.atol,
                                 self.rtol, self.max_iter)
        return u1

    def step_bte(self, v):
        """:2030-2071: x half step, v step under bs_E, x half step."""
        p = self.param
        v = self.xstep.step(v)
        vl = self.ords_to_sph(v)
        vl = kr.vstep_exact(self.K0, self.A, self.bs_E, self.dt, p.tau, vl)
        v = self.sph_to_ords(vl)
        return self.xstep.step(v)

    def step_init(self, u, v):
        """:2073-2212 (the part that acts on the state): bte_to_fluid and fluid_to_bte at t = 0."""
        u = self.bte_to_fluid(u, v)
        self.fluid_to_bte(u, 0.0)
        return u

    def step(self, u, v, tt):
        """:2214-2241, Strang splitting with the frozen field."""
        dt = self.dt
        E = self.field(u, tt)
        u = self.step_fluid(u, E, 0.5 * dt)
        self.fluid_to_bte(u, tt)
        v = self.step_bte(v)
        u = self.bte_to_fluid(u, v)
        E = self.field(u, tt + 0.5 * dt)
        u = self.step_fluid(u, E, 0.5 * dt)
        return u, v
