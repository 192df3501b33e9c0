"""Oracle: Galerkin operators of the B-spline x spherical-harmonic basis (FP64).

Ordering is radial-major, index = p * num_sh + l (spec_spherical.py:273-275).
Restates (BESolver/python/...):
  * mass matrix, spline branch           spec_spherical.py:226-277 (10 pts/interval)
  * Cholesky inverse of the mass matrix   spec_spherical.py:299-336
  * E-field advection, spline branch      spec_spherical.py:442-498 (10 pts/interval)
    with AM/BM/AD/BD                      advection_operator_spherical_polys.py:6-16
  * binary collision operator, CG branch  collision_operator_spherical.py:96-225
  * gas-temperature operator, CG branch   collision_operator_spherical.py:443-497
  * QoI vectors mass/temp/mobility/diffusion/rates  utils.py:73-150, 321-395, 854-944
  * Maxwellian projection function_to_basis  utils.py:475-532 (isotropic case)
The reference evaluates each (p,k) entry as a masked dot product over the
quadrature points in supp(B_k) (or supp(B_p)); B-splines vanish outside their
support, so the same sums are written here as products of basis tables over
all points, which adds only exact zeros.

Pinned by tests/test_oracle_operators.py: exact polynomial integrals (sympy),
partition-of-unity identities (M^-1 m = 1, column sums of the l=0 collision
block vanish for elastic/excitation and equal the rate vector for ionization),
scipy.integrate.quad on individual entries, and Maxwellian closed forms for
the QoI vectors.
"""
import math
import numpy as np
import scipy.constants as sc
import scipy.linalg

from . import bspline
from .xsec import interp_linear_flat

QE = sc.e
ME = sc.electron_mass
KB = sc.Boltzmann
EV = sc.electron_volt
EPS0 = sc.epsilon_0
TEMP_K_1EV = EV / KB
C_GAMMA = math.sqrt(2.0 * QE / ME)


def vth_from_Te(Te_ev):
    """collisions.py:20-25 with T = Te * TEMP_K_1EV."""
    return np.sqrt(2.0 * KB * (Te_ev * TEMP_K_1EV) / ME)


def maxwellian0(vth):
    """utils.py:713-715 get_maxwellian_3d(vth,1)(0) = 1/(sqrt(pi) vth)^3."""
    return 1.0 / (math.sqrt(math.pi) * vth) ** 3


def sph_norm(l):
    return math.sqrt((2.0 * l + 1) / (4.0 * math.pi))


def AM(l):
    return l / (2.0 * l + 1.0) * sph_norm(l) / sph_norm(l - 1)


def BM(l):
    return (l + 1.0) / (2.0 * l + 1.0) * sph_norm(l) / sph_norm(l + 1)


def AD(l):
    return l * (l + 1.0) / (2.0 * l + 1.0) * sph_norm(l) / sph_norm(l - 1)


def BD(l):
    return -l * (l + 1.0) / (2.0 * l + 1.0) * sph_norm(l) / sph_norm(l + 1)


class Grid:
    """One velocity-space grid: knot vector, radial basis count, SH modes, quadrature order.

    ``num_p`` = Nr + 1 radial functions (bte_0d3v_batched.py:186); ``lmax`` gives
    m = 0 modes l = 0..lmax; ``q`` is spline_qpts (points per knot interval used
    by the collision operators, spec_sp._num_q_radial / num_knot_intervals)."""

    def __init__(self, t, num_p, lmax, q, vth, sp_order=3):
        self.t = np.asarray(t, dtype=np.float64)
        self.num_p = int(num_p)
        self.lmax = int(lmax)
        self.num_sh = self.lmax + 1
        self.q = int(q)
        self.vth = float(vth)
        self.sp_order = sp_order
        self.n = self.num_p * self.num_sh
        assert len(self.t) == self.num_p + sp_order + 1

    def quad(self, per_interval):
        return bspline.gauss_pn(self.t, per_interval)

    def V(self, x, dorder=0):
        return bspline.vandermonde(self.t, self.sp_order, self.num_p, x, dorder)

    def scatter_radial(self, blocks):
        """blocks: dict l -> (num_p,num_p) radial matrix; returns (n,n)."""
        A = np.zeros((self.n, self.n))
        for l, B in blocks.items():
            A[l::self.num_sh, l::self.num_sh] = B
        return A


def grid_single(Nr, lmax, Te_ev, ev_max, q, sp_order=3):
    """Single 0D path grid: k_domain=(0, sqrt(ev_max) c_gamma / vth), no extension and
    no knot snapping (bte_0d3v_solver.py:161-190 with use_dg = 0)."""
    vth = float(vth_from_Te(Te_ev))
    kd = (0.0, math.sqrt(ev_max) * C_GAMMA / vth)
    t, _ = bspline.make_knots(kd, sp_order, Nr + 1, sig_pts=None, extend_domain=False)
    return Grid(t, Nr + 1, lmax, q, vth, sp_order)


def grid_batched(Nr, lmax, Te_ev, ev_max, q, thresholds, sp_order=3, extend=True, extend_log=False):
    """Batched 0D path grid (bte_0d3v_batched.py:310-349): thresholds snap knots
    (dg_nodes = sqrt(eps_th) c_gamma / vth) and the LINEAR extension scales
    the upper end by 1.2 (vspace_grid_extension.LINEAR, the default); extend_log (with
    extend=False) is the LOG extension (basis.py:347-349)."""
    vth = float(vth_from_Te(Te_ev))
    kd = (0.0, math.sqrt(ev_max) * C_GAMMA / vth)
    sig = np.sort(np.array(sorted(set([th for th in thresholds if th is not None and th > 0]))))
    dg_nodes = np.sqrt(sig) * C_GAMMA / vth
    t, _ = bspline.make_knots(kd, sp_order, Nr + 1, sig_pts=dg_nodes, extend_domain=extend,
                              extend_domain_with_log=extend_log)
    return Grid(t, Nr + 1, lmax, q, vth, sp_order)


# ----------------------------------------------------------------------------- M, A

def mass_matrix(g: Grid):
    """spec_spherical.py:226-277: M[(p,l),(k,l)] = sum w x^2 B_p B_k, 2*(p+2) pts/interval."""
    x, w = g.quad(2 * (g.sp_order + 2))
    V = g.V(x)
    mm = (V * (w * x * x)) @ V.T
    # the reference only fills |p-k| < sp_order+3; entries beyond are exact zeros anyway
    return g.scatter_radial({l: mm for l in range(g.num_sh)})


def inverse_mass_matrix(M):
    """spec_spherical.py:299-336: Cholesky, Linv by triangular solve, Minv = Linv^T Linv."""
    L = np.linalg.cholesky(M)
    Linv = scipy.linalg.solve_triangular(L, np.identity(M.shape[0]), lower=True)
    return Linv.T @ Linv


def advection_radial(g: Grid):
    """mm1[p,k] = sum w x^2 B_p B_k', mm2[p,k] = sum w x B_p B_k (spec_spherical.py:454-474)."""
    x, w = g.quad((g.sp_order + 2) * 2)
    V = g.V(x)
    D = g.V(x, 1)
    mm1 = (V * (w * x * x)) @ D.T
    mm2 = (V * (w * x)) @ V.T
    return mm1, mm2


def advection_matrix(g: Grid):
    """spec_spherical.py:482-498: couple (p,q) to (k,l) for l = q +- 1."""
    mm1, mm2 = advection_radial(g)
    ns = g.num_sh
    A = np.zeros((g.num_p, ns, g.num_p, ns))
    for qi in range(ns):
        l = qi + 1
        if l < ns:
            A[:, qi, :, l] = mm1 * AM(l) + AD(l) * mm2
        l = qi - 1
        if l >= 0:
            A[:, qi, :, l] = mm1 * BM(l) + BD(l) * mm2
    return A.reshape(g.n, g.n)


# ----------------------------------------------------------------------------- collisions

class Process:
    """One electron-heavy process (collisions.py:412-431 + load_cross_section)."""

    def __init__(self, ptype, energy, sigma, threshold=0.0, mass_ratio=1.36e-5,
                 attachment_energy=0.0, species="Ar", name=""):
        self.type = ptype.upper()
        self.energy = np.asarray(energy, dtype=np.float64)
        self.sigma_tab = np.asarray(sigma, dtype=np.float64)
        self.threshold = float(threshold)
        self.mByM = float(mass_ratio)
        self.attachment_energy = float(attachment_energy)
        self.species = species
        self.name = name

    def total_cross_section(self, ev):
        return interp_linear_flat(self.energy, self.sigma_tab, ev)


def processes_from_lxcat(recs):
    """Mirror of bte_0d3v_batched.py:186-214: mByM of every process is the elastic m/M."""
    mbym = None
    for r in recs:
        if r["type"] == "ELASTIC":
            mbym = r["mass_ratio"]
            break
    return [Process(r["type"], r["energy"], r["cross section"], r["threshold"], mbym,
                    r["attachment_energy"], r["species"], r["process"]) for r in recs]


def elastic_gain_difference(g: Grid, x, c_mu):
    """B_p(x sqrt(1 - c_mu)) - B_p(x) for every basis function p (rows) and point x (columns),
    as minus the integral of B_p' over [x - dv, x], dv = x c_mu / (1 + sqrt(1 - c_mu)), split at
    the knot inside (dv is far below the knot spacing), 2 Gauss-Legendre points per piece
    (B_p' is quadratic on a piece).  Piece widths come from dv and from x minus the knot, not
    from the rounded end point x - dv."""
    x = np.asarray(x, dtype=np.float64)
    dv = x * c_mu / (1.0 + math.sqrt(1.0 - c_mu))
    t = g.t
    kn = np.searchsorted(t, x - dv, side="right")      # first knot > x - dv
    knot = t[np.minimum(kn, len(t) - 1)]
    inside = (kn < len(t)) & (knot < x)
    roff = np.where(inside, x - knot, 0.0)             # the split point is x - roff
    out = np.zeros((g.num_p, x.size))
    c = 1.0 / (2.0 * math.sqrt(3.0))
    for a_off, b_off in ((dv, roff), (roff, np.zeros_like(x))):
        h = a_off - b_off
        m = x - 0.5 * (a_off + b_off)
        for y in (m - c * h, m + c * h):
            out -= 0.5 * h * g.V(y, 1)
    return out


def collision_radial(g: Grid, proc: Process, tgK=0.0):
    """(q0, q1) radial blocks of _LOp_eulerian_radial_only, CG branch
    (collision_operator_spherical.py:96-223).  q0 is the l = 0 block, q1 the l >= 1 block."""
    vth = g.vth
    x, w = g.quad(g.q)
    ev = (x * vth / C_GAMMA) ** 2
    if proc.type == "ELASTIC":
        sig = proc.total_cross_section(ev)
        c_mu = 2.0 * proc.mByM
        vpost = x * math.sqrt(1.0 - c_mu)
        kappa = (KB * tgK * c_mu * 0.5 / ME) / vth
        V = g.V(x)
        D = g.V(x, 1)
        wt = w * vth * x ** 3 * sig
        # B_p(v') - B_p(v) (collision_operator_spherical.py:136-144) evaluated as the integral of
        # B_p' over [v', v] (2 Gauss points per knot piece, exact): the same quantity without the
        # cancellation of two values that differ by ~m/M (which leaves ~1e-11 relative noise)
        dV = elastic_gain_difference(g, x, c_mu)
        q0 = dV @ (V * wt).T - kappa * (D * (w * x ** 3 * sig)) @ D.T
        q1 = -(V * wt) @ V.T
        return q0, q1
    if proc.type in ("EXCITATION", "IONIZATION", "ATTACHMENT"):
        split = 2.0 if proc.type == "IONIZATION" else 1.0
        if proc.type in ("EXCITATION", "IONIZATION"):
            keep = ev >= proc.threshold                     # collision_operator_spherical.py:153-155
            x, w, ev = x[keep], w[keep], ev[keep]
            vpost = C_GAMMA * np.sqrt((1.0 / split) * (ev - proc.threshold)) / vth
        else:
            vpost = C_GAMMA * np.sqrt((1.0 / split) * (ev + proc.attachment_energy)) / vth
        sig = proc.total_cross_section(ev)
        V = g.V(x)
        Vp = g.V(vpost)
        wt = w * vth * x ** 3 * sig
        q0 = (split * Vp - V) @ (V * wt).T
        q1 = -(V * wt) @ V.T
        return q0, q1
    raise NotImplementedError(proc.type)


def collision_matrix(g: Grid, proc: Process, tgK=0.0):
    q0, q1 = collision_radial(g, proc, tgK)
    blocks = {0: q0}
    for l in range(1, g.num_sh):
        blocks[l] = q1
    return g.scatter_radial(blocks)


def gas_temperature_radial(g: Grid, proc: Process):
    """electron_gas_temperature, CG branch (collision_operator_spherical.py:463-497), tgK = 1."""
    assert proc.type == "ELASTIC"
    vth = g.vth
    x, w = g.quad(g.q)
    sig = proc.total_cross_section((x * vth / C_GAMMA) ** 2)
    c_mu = 2.0 * proc.mByM
    kappa = (KB * 1.0 * c_mu * 0.5 / ME) / vth
    D = g.V(x, 1)
    return -kappa * (D * (w * x ** 3 * sig)) @ D.T


def gas_temperature_matrix(g: Grid, proc: Process):
    return g.scatter_radial({0: gas_temperature_radial(g, proc)})


# ----------------------------------------------------------------------------- QoI vectors

def mass_op(g: Grid):
    """utils.py:73-110 (spline): sqrt(4 pi) * sum w x^2 B_k on l = 0, 10 pts/interval."""
    x, w = g.quad(2 * (g.sp_order + 2))
    v = math.sqrt(4.0 * math.pi) * (g.V(x) @ (w * x ** 2))
    out = np.zeros(g.n)
    out[0::g.num_sh] = v
    return out


def temp_op(g: Grid):
    """utils.py:321-350 (spline): sqrt(4 pi) * sum w x^4 B_k on l = 0, 12 pts/interval."""
    x, w = g.quad(2 * (g.sp_order + 3))
    v = math.sqrt(4.0 * math.pi) * (g.V(x) @ (w * x ** 4))
    out = np.zeros(g.n)
    out[0::g.num_sh] = v
    return out


TEMP_EV_FAC = 0.5 * ME * (2.0 / 3.0 / KB) / TEMP_K_1EV   # bte_0d3v_batched.py:394


def mobility_op(g: Grid):
    """utils.py:890-913: (2 vth^4 / c^4) sum w x^3 B_p, 12 pts/interval."""
    x, w = g.quad(2 * (g.sp_order + 3))
    return (2 * g.vth ** 4 / C_GAMMA ** 4) * (g.V(x) @ (w * x ** 3))


def diffusion_op(g: Grid, procs):
    """utils.py:915-944: (2 vth^4 / c^4) sum w x^3 B_p / sigma_total, q pts/interval."""
    x, w = g.quad(g.q)
    ev = (x * g.vth / C_GAMMA) ** 2
    tot = sum(p.total_cross_section(ev) for p in procs)
    return (2 * g.vth ** 4 / C_GAMMA ** 4) * (g.V(x) @ (w * x ** 3 / tot))


def rate_op(g: Grid, proc):
    """utils.py:854-888: (2 vth^4 / c^3) sum w x^3 sigma_g B_p, q pts/interval."""
    x, w = g.quad(g.q)
    ev = (x * g.vth / C_GAMMA) ** 2
    return (2 * g.vth ** 4 / C_GAMMA ** 3) * (g.V(x) @ (w * x ** 3 * proc.total_cross_section(ev)))


def maxwellian_init(g: Grid, Minv):
    """bte_0d3v_batched.py:499-502 / utils.py:475-532 for hv = exp(-v^2)/pi^1.5.

    The angular quadrature of function_to_basis integrates Y_00 to sqrt(4 pi)
    and the l >= 1 modes to zero (up to rounding) for an isotropic function."""
    x, w = g.quad(g.q)
    hq = np.exp(-x ** 2) / math.pi ** 1.5 * x ** 2
    b = np.zeros(g.n)
    b[0::g.num_sh] = (4.0 * math.pi * (1.0 / (2.0 * math.sqrt(math.pi)))) * (g.V(x) @ (w * hq))
    return Minv @ b
