"""The reference's radial basis object, basis.BSpline (basis.py:327-504), over this library.

Same constructor arguments and attributes the assembly reads (_t, _sp_order, _num_p,
_domain, _kdomain, _t_unique, _num_knot_intervals, _dg_idx) and the same knot rules,
including the LINEAR (1.2x) and LOG (5x, geometric) domain extensions and the snapping of
significant points.  Basis values are evaluated by the HIP kernel behind bz_basis_table
and returned as host arrays, as the reference's Pn / diff do.  Discontinuous (DG) splines
are outside this library's scope and raise NotImplementedError.
"""
import enum

import numpy as np
import torch

from . import _lib
from .vspace import make_knots


class BasisType(enum.Enum):
    SPLINES = 0


class BSpline:
    def __init__(self, k_domain, spline_order, num_p, q_per_knot=None, sig_pts=list(), knots_vec=None,
                 dg_splines=False, verbose=True, extend_domain=False, extend_domain_with_log=False):
        if dg_splines:
            raise NotImplementedError("discontinuous splines (dg_splines) are not part of this library")
        if spline_order != 3:
            raise NotImplementedError("the HIP kernels implement cubic B-splines (spline_order 3)")
        self._basis_type = BasisType.SPLINES
        self._kdomain = (k_domain[0], k_domain[1])
        self._window = k_domain
        self._sp_order = spline_order
        self._sig_pts = sig_pts
        if knots_vec is None:
            self._t = make_knots(k_domain, spline_order, num_p, sig_pts, extend_domain=extend_domain,
                                 extend_domain_with_log=extend_domain_with_log)
            if extend_domain:
                self._domain = (k_domain[0], 1.2 * k_domain[1])
            elif extend_domain_with_log:
                self._domain = (k_domain[0], 5 * k_domain[1])
            else:
                self._domain = tuple(k_domain)
            self._num_p = num_p
        else:
            self._t = np.asarray(knots_vec, dtype=np.float64)
            self._domain = tuple(k_domain)
            self._kdomain = (self._t[0], self._t[-1])
            self._num_p = len(self._t) - (spline_order + 1)
        self._ele = 1
        self._ele_p = self._num_p
        self._t_unique = np.unique(self._t)
        self._num_knot_intervals = len(self._t_unique) - 1
        self._dg_idx = [0, self._num_p - 1]
        self._dev = _lib.require_device()
        self._t_dev = torch.from_numpy(np.ascontiguousarray(self._t)).to(self._dev)

    def _table(self, x, dorder):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
        xd = torch.from_numpy(x).to(self._dev)
        out = torch.empty((self._num_p, len(x)), dtype=torch.float64, device=self._dev)
        _lib.call("bz_basis_table", self._t_dev, self._num_p, xd, len(x), int(dorder), out)
        return out.cpu().numpy()

    def Pn(self, deg):
        """B_deg as a function of (x, l); zero outside its support (basis.py:411-412)."""
        return lambda x, l: self._table(x, 0)[deg].reshape(np.shape(x))

    def diff(self, deg, dorder):
        """d^dorder B_deg / dx^dorder as a function of (x, l) (basis.py:499-504)."""
        if dorder not in (0, 1):
            raise NotImplementedError("bz_basis_table evaluates values and first derivatives")
        return lambda x, l: self._table(x, dorder)[deg].reshape(np.shape(x))

    def Gauss_Pn(self, deg):
        """deg // (number of knot intervals) Gauss-Legendre points on every non-empty knot
        interval (basis.py:414-429)."""
        q = deg // self._num_knot_intervals
        x, w = np.polynomial.legendre.leggauss(q)
        qx, qw = [], []
        t = self._t
        for i in range(len(t) - 1):
            if t[i] != t[i + 1]:
                qx.append(0.5 * (t[i + 1] - t[i]) * x + 0.5 * (t[i + 1] + t[i]))
                qw.append(0.5 * (t[i + 1] - t[i]) * w)
        qx, qw = np.concatenate(qx), np.concatenate(qw)
        assert np.allclose(np.sum(qw), t[-1] - t[0], rtol=1e-14, atol=1e-14)
        return qx, qw
