"""Device-resident velocity-space grid and Galerkin assembly (HIP kernels via libbz.so).

A ``VGrid`` holds the cubic B-spline knot vector, the spherical-harmonic modes
(m = 0), the thermal velocity and the quadrature tables on the GPU, and exposes
the operator assembly of the reference (all results are FP64 torch tensors on
the current HIP device, radial-major index p*ns + l):

  mass_matrix / inverse_mass_matrix     spec_spherical.py:226-336
  advection_matrix                      spec_spherical.py:442-498
  collision_matrix (elastic / excitation / ionization / attachment)
                                        collision_operator_spherical.py:79-225
  gas_temperature_matrix                collision_operator_spherical.py:443-578
  QoI vectors mass/temp/mobility/diffusion/rates   utils.py:73-150, 321-395, 854-944

Host work is limited to the knot vector and the Gauss-Legendre reference
nodes (tables of constants) and to launching kernels.
"""
import math

import numpy as np
import scipy.constants as sc
import torch

from . import _lib

QE = sc.e
ME = sc.electron_mass
KB = sc.Boltzmann
EV = sc.electron_volt
TEMP_K_1EV = EV / KB
C_GAMMA = math.sqrt(2.0 * QE / ME)
TEMP_EV_FAC = 0.5 * ME * (2.0 / 3.0 / KB) / TEMP_K_1EV

PTYPE = {"ELASTIC": 0, "EFFECTIVE": 0, "EXCITATION": 1, "IONIZATION": 2, "ATTACHMENT": 3}


def vth_from_Te(Te_ev):
    """collisions.electron_thermal_velocity(Te * TEMP_K_1EV) (collisions.py:20-25)."""
    return float(np.sqrt(2.0 * KB * (Te_ev * TEMP_K_1EV) / ME))


def uniform_knots(k_domain, num_p, sp_order):
    """basis.BSpline.uniform_knots (basis.py:639-653)."""
    t = np.full(sp_order, float(k_domain[0]))
    t = np.append(t, np.linspace(k_domain[0], k_domain[1], num_p - sp_order, endpoint=False))
    return np.append(t, np.full(sp_order + 1, float(k_domain[1])))


def log_extended_knots(k_domain, num_p, sp_order, ext_kdomain):
    """basis.BSpline.uniform_knots_with_extended_bdy (basis.py:680-711): the lower half of
    the knots uniform on the domain, the rest in a geometric progression of base 1.2 from
    its end to ext_kdomain."""
    n1 = int(math.ceil(0.5 * num_p))
    n2 = num_p - n1
    head = np.full(sp_order, float(k_domain[0]))
    lin = np.linspace(k_domain[0], k_domain[1], n1, endpoint=False)
    b = 1.2
    geo = np.logspace(math.log(k_domain[1]) / math.log(b), math.log(ext_kdomain) / math.log(b), n2 - sp_order,
                      base=b, endpoint=False)
    return np.concatenate([head, lin, geo, np.full(sp_order + 1, float(ext_kdomain))])


def make_knots(k_domain, sp_order, num_p, sig_pts=None, extend_domain=False, extend_domain_with_log=False):
    """basis.BSpline.__init__ knot rules (basis.py:338-362): optional 1.2x extension or the
    LOG extension to 5x the domain, and snapping of the nearest interior knot to each
    significant point."""
    if extend_domain:
        t = uniform_knots((k_domain[0], 1.2 * k_domain[1]), num_p, sp_order)
    elif extend_domain_with_log:
        t = log_extended_knots(k_domain, num_p, sp_order, 5 * k_domain[1])
    else:
        t = uniform_knots(k_domain, num_p, sp_order)
    if sig_pts is not None:
        for s in np.asarray(sig_pts, dtype=np.float64).reshape(-1):
            if k_domain[0] < s < k_domain[1]:
                idx = int(np.argmin(np.abs(t - s)))
                if sp_order + 1 < idx < len(t) - sp_order - 1:
                    t[idx] = s
    return t


class XSection:
    """One electron-heavy process: type, threshold, m/M and its cross-section table on device."""

    def __init__(self, ptype, energy, sigma, threshold=0.0, mass_ratio=1.36e-5, attachment_energy=0.0,
                 species="", name="", device=None):
        self.type = ptype.upper()
        self.ptype = PTYPE[self.type]
        self.energy_np = np.ascontiguousarray(energy, dtype=np.float64)
        self.sigma_np = np.ascontiguousarray(sigma, dtype=np.float64)
        self.threshold = float(threshold)
        self.mByM = float(mass_ratio)
        self.attachment_energy = float(attachment_energy)
        self.species = species
        self.name = name
        dev = device if device is not None else _lib.require_device()
        self.energy = torch.from_numpy(self.energy_np).to(dev)
        self.sigma = torch.from_numpy(self.sigma_np).to(dev)


class VGrid:
    """One velocity grid (np_r = Nr+1 cubic B-splines x (lmax+1) m=0 harmonics)."""

    def __init__(self, t, num_p, lmax, q, vth, sp_order=3, device=None):
        assert sp_order == 3, "the HIP kernels implement cubic B-splines"
        self.dev = device if device is not None else _lib.require_device()
        self.t_np = np.ascontiguousarray(t, dtype=np.float64)
        self.num_p = int(num_p)
        self.npr = self.num_p
        self.lmax = int(lmax)
        self.num_sh = self.lmax + 1
        self.n = self.num_p * self.num_sh
        self.q = int(q)
        self.vth = float(vth)
        self.sp_order = sp_order
        assert len(self.t_np) == self.num_p + sp_order + 1
        self.t = torch.from_numpy(self.t_np).to(self.dev)
        self._gl = {}
        self.num_knot_intervals = len(np.unique(self.t_np)) - 1

    # -- tables
    def gl(self, q):
        if q not in self._gl:
            x, w = np.polynomial.legendre.leggauss(q)
            self._gl[q] = (torch.from_numpy(x).to(self.dev), torch.from_numpy(w).to(self.dev))
        return self._gl[q]

    def _empty(self, *shape):
        return torch.empty(shape, dtype=torch.float64, device=self.dev)

    # -- radial integrals
    def radial_bilinear(self, q, pw, da, db):
        glx, glw = self.gl(q)
        out = self._empty(self.npr, self.npr)
        _lib.call("bz_radial_bilinear", self.t, self.npr, glx, glw, q, pw, da, db, out)
        return out

    def radial_moment(self, q, pw, mode=0, scale=1.0, xsecs=()):
        glx, glw = self.gl(q)
        out = self._empty(self.npr)
        if mode in (0, 3):
            e = s = off = None
            ntab = 0
            dummy = self._empty(1)
            e = s = dummy
            off = torch.zeros(2, dtype=torch.int32, device=self.dev)
        else:
            e = torch.cat([x.energy for x in xsecs]).contiguous()
            s = torch.cat([x.sigma for x in xsecs]).contiguous()
            lens = [0] + [len(x.energy_np) for x in xsecs]
            off = torch.tensor(np.cumsum(lens), dtype=torch.int32, device=self.dev)
            ntab = len(xsecs)
        _lib.call("bz_radial_moment", self.t, self.npr, glx, glw, q, pw, mode, float(scale), self.vth, C_GAMMA,
                  e, s, off, ntab, out)
        return out

    def scatter_diag(self, blocks, lmap, scale=1.0):
        """blocks: (nb, npr, npr) tensor; lmap[l] -> block index or -1."""
        blk = blocks.contiguous()
        lm = torch.tensor(lmap, dtype=torch.int32, device=self.dev)
        out = self._empty(self.n, self.n)
        _lib.call("bz_scatter_diag_blocks", blk, self.npr, self.num_sh, lm, float(scale), out)
        return out

    # -- operators
    def mass_radial(self):
        return self.radial_bilinear(2 * (self.sp_order + 2), 2, 0, 0)

    def mass_matrix(self):
        mm = self.mass_radial()
        return self.scatter_diag(mm.reshape(1, self.npr, self.npr), [0] * self.num_sh)

    def mass_radial_inverse(self, mm=None):
        mm = self.mass_radial() if mm is None else mm
        lb = self._empty(self.npr, 4)
        _lib.call("bz_band_cholesky", mm, self.npr, 3, lb)
        inv = self._empty(self.npr, self.npr)
        work = self._empty(self.npr, self.npr)
        _lib.call("bz_band_chol_inverse", lb, self.npr, 3, inv, work)
        return inv

    def inverse_mass_matrix(self):
        inv = self.mass_radial_inverse()
        return self.scatter_diag(inv.reshape(1, self.npr, self.npr), [0] * self.num_sh)

    def advection_radial(self):
        q = (self.sp_order + 2) * 2
        return self.radial_bilinear(q, 2, 0, 1), self.radial_bilinear(q, 1, 0, 0)

    def advection_matrix(self, scale=1.0):
        mm1, mm2 = self.advection_radial()
        ns = self.num_sh
        cup = np.zeros(2 * ns)
        cdn = np.zeros(2 * ns)
        N = lambda l: math.sqrt((2.0 * l + 1) / (4.0 * math.pi))
        for qi in range(ns):
            l = qi + 1
            if l < ns:   # AM, AD of (l, 0)
                cup[2 * qi] = l / (2.0 * l + 1.0) * N(l) / N(l - 1)
                cup[2 * qi + 1] = l * (l + 1.0) / (2.0 * l + 1.0) * N(l) / N(l - 1)
            l = qi - 1
            if l >= 0:   # BM, BD of (l, 0)
                cdn[2 * qi] = (l + 1.0) / (2.0 * l + 1.0) * N(l) / N(l + 1)
                cdn[2 * qi + 1] = -l * (l + 1.0) / (2.0 * l + 1.0) * N(l) / N(l + 1)
        cu = torch.from_numpy(cup).to(self.dev)
        cd = torch.from_numpy(cdn).to(self.dev)
        out = self._empty(self.n, self.n)
        _lib.call("bz_advection_scatter", mm1, mm2, self.npr, ns, cu, cd, float(scale), out)
        return out

    def collision_radial(self, xs: XSection, tgK=0.0):
        glx, glw = self.gl(self.q)
        q0 = self._empty(self.npr, self.npr)
        q1 = self._empty(self.npr, self.npr)
        c_mu = 2.0 * xs.mByM
        kappa = (KB * tgK * c_mu * 0.5 / ME) / self.vth if xs.ptype == 0 else 0.0
        _lib.call("bz_collision_radial", self.t, self.npr, glx, glw, self.q, xs.ptype, self.vth, C_GAMMA, xs.mByM,
                  xs.threshold, xs.attachment_energy, kappa, xs.energy, xs.sigma, len(xs.energy_np), q0, q1)
        return q0, q1

    def collision_matrix(self, xs: XSection, tgK=0.0, scale=1.0):
        q0, q1 = self.collision_radial(xs, tgK)
        blocks = torch.stack([q0, q1])
        return self.scatter_diag(blocks, [0] + [1] * (self.num_sh - 1), scale)

    def gas_temperature_radial(self, xs: XSection):
        assert xs.ptype == 0
        glx, glw = self.gl(self.q)
        out = self._empty(self.npr, self.npr)
        kappa1 = (KB * 1.0 * 2.0 * xs.mByM * 0.5 / ME) / self.vth
        _lib.call("bz_gas_temperature_radial", self.t, self.npr, glx, glw, self.q, self.vth, C_GAMMA, kappa1,
                  xs.energy, xs.sigma, len(xs.energy_np), out)
        return out

    def gas_temperature_matrix(self, xs: XSection, scale=1.0):
        r = self.gas_temperature_radial(xs)
        return self.scatter_diag(r.reshape(1, self.npr, self.npr), [0] + [-1] * (self.num_sh - 1), scale)

    # -- QoI vectors (utils.py)
    def _l0(self, v):
        out = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        out[0::self.num_sh] = v
        return out

    def mass_op(self):
        return self._l0(self.radial_moment(2 * (self.sp_order + 2), 2, 0, math.sqrt(4.0 * math.pi)))

    def temp_op(self):
        return self._l0(self.radial_moment(2 * (self.sp_order + 3), 4, 0, math.sqrt(4.0 * math.pi)))

    def mobility_op(self):
        return self.radial_moment(2 * (self.sp_order + 3), 3, 0, 2 * self.vth ** 4 / C_GAMMA ** 4)

    def diffusion_op(self, xsecs):
        return self.radial_moment(self.q, 3, 2, 2 * self.vth ** 4 / C_GAMMA ** 4, xsecs)

    def rate_op(self, xs):
        return self.radial_moment(self.q, 3, 1, 2 * self.vth ** 4 / C_GAMMA ** 3, [xs])

    def maxwellian_init(self):
        """Maxwellian h0 = M^-1 [sqrt(4 pi) pi^-1.5 sum_x w x^2 exp(-x^2) B_k] on l = 0
        (bte_0d3v_batched.initialize, bte_0d3v_batched.py:481-503, via function_to_basis,
        utils.py:475-532; its angular quadrature leaves only the l = 0 mode)."""
        rhs = self.radial_moment(self.q, 2, 3, math.sqrt(4.0 * math.pi) / math.pi ** 1.5)
        inv = self.mass_radial_inverse()
        h0 = self._empty(self.npr)
        _lib.call("bz_gemv", inv, self.npr, self.npr, 0, rhs, 1.0, h0)
        return self._l0(h0)
