"""The reference's spectral expansion object, spec_spherical.SpectralExpansionSpherical
(spec_spherical.py:20-498), for the B-spline x real spherical harmonic (m = 0) basis.

Same constructor and the methods the 0D assembly calls (bte_0d3v_batched.py:351-372):
coefficient counts, create_vec / create_mat, compute_mass_matrix, inverse_mass_mat,
Vq_r / Vdq_r, compute_advection_matix and the radial basis evaluation.  The matrices are
assembled by this library's HIP kernels and returned as torch tensors on the HIP device
(the reference assembles numpy arrays on the host and copies them to the device in
host_to_device_setup).  Only cubic B-spline radial bases and m = 0 harmonics l = 0..lmax
are supported; the polynomial radial bases, hemispherical harmonics and the ordinate / FVM
discretisations of the reference are outside this library.
"""
import numpy as np
import torch

from .vspace import VGrid


class SpectralExpansionSpherical:
    def __init__(self, p_order, basis_p, sph_harm_lm, domain=None, window=None):
        ls = [l for l, _ in sph_harm_lm]
        if any(m != 0 for _, m in sph_harm_lm) or ls != list(range(len(ls))):
            raise NotImplementedError("harmonics (l, 0) for l = 0..lmax in order are supported")
        self._sph_harm_lm = list(sph_harm_lm)
        self._p = int(p_order)
        self._basis_p = basis_p
        self._domain = domain
        self._window = window
        self._num_q_radial = None
        self._r_grid = [window]
        self._grids = {}
        assert basis_p._num_p == self._p + 1

    # ------------------------------------------------------------------ sizes
    def get_radial_basis(self):
        return self._basis_p

    def get_radial_basis_type(self):
        return self._basis_p._basis_type

    def get_num_coefficients(self):
        return (self._p + 1) * len(self._sph_harm_lm) * len(self._r_grid)

    def get_num_radial_domains(self):
        return len(self._r_grid)

    def get_dof_per_radial_domain(self):
        return (self._p + 1) * len(self._sph_harm_lm)

    def create_vec(self, dtype=float):
        return np.zeros((self.get_num_coefficients(), 1), dtype=dtype)

    def create_mat(self, dtype=float):
        n = self.get_num_coefficients()
        return np.zeros((n, n), dtype=dtype)

    # ------------------------------------------------------------------ basis
    def basis_eval_radial(self, r, k, l):
        return self._basis_p.Pn(k)(r, l)

    def basis_derivative_eval_radial(self, r, k, l, dorder):
        return self._basis_p.diff(k, dorder)(r, l)

    def Vq_r(self, v_r, l, scale=1):
        """(num_p,) + v_r.shape basis Vandermonde (spec_spherical.py:338-352)."""
        v = np.asarray(v_r, dtype=np.float64)
        return scale * self._basis_p._table(v, 0).reshape((self._p + 1,) + v.shape)

    def Vdq_r(self, v_r, l, d_order=1, scale=1):
        """Derivative Vandermonde (spec_spherical.py:369-383); d_order 1."""
        if d_order != 1:
            raise NotImplementedError("first derivatives only")
        v = np.asarray(v_r, dtype=np.float64)
        return scale * self._basis_p._table(v, 1).reshape((self._p + 1,) + v.shape)

    # ------------------------------------------------------------------ operators
    def vgrid(self, vth=1.0):
        """The device grid behind this expansion (quadrature: _num_q_radial points over the
        knot intervals, the reference's spline_qpts per interval)."""
        nki = self._basis_p._num_knot_intervals
        q = (self._num_q_radial // nki) if self._num_q_radial else 5
        key = (q, float(vth))
        if key not in self._grids:
            self._grids[key] = VGrid(self._basis_p._t, self._p + 1, len(self._sph_harm_lm) - 1, q, vth)
        return self._grids[key]

    def compute_mass_matrix(self, v_th=1.0):
        """M[(p,l),(k,l)] = int v^2 B_p B_k (spec_spherical.py:193-297), (n, n) on the device."""
        return self.vgrid().mass_matrix()

    def inverse_mass_mat(self, v_th=1, Mmat=None):
        """Block inverse of the mass matrix through its banded Cholesky factor
        (spec_spherical.py:299-336)."""
        g = self.vgrid()
        if Mmat is None:
            return g.inverse_mass_matrix()
        ns = len(self._sph_harm_lm)
        mm = torch.as_tensor(Mmat, dtype=torch.float64, device=g.dev)[0::ns, 0::ns].contiguous()
        inv = g.mass_radial_inverse(mm)
        return g.scatter_diag(inv.reshape(1, g.npr, g.npr), [0] * ns)

    def compute_advection_matix(self):
        """E-field advection operator (spec_spherical.py:432-498), (n, n) on the device, in
        the reference's units (the 0D driver scales it by (1 / vth) (q / m_e))."""
        return self.vgrid().advection_matrix()

    def gl_vr(self, Nvr, use_bspline_qgrid=False):
        """Nvr Gauss-Legendre points on the basis domain (spec_spherical.py:1010-1018), or Nvr
        points per knot interval with use_bspline_qgrid."""
        if use_bspline_qgrid:
            return self._basis_p.Gauss_Pn(Nvr * self._basis_p._num_knot_intervals)
        x, w = np.polynomial.legendre.leggauss(Nvr)
        a, b = self._basis_p._domain
        return 0.5 * (b - a) * x + 0.5 * (b + a), 0.5 * (b - a) * w

