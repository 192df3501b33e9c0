"""Oracle: clamped B-spline knots, per-interval Gauss-Legendre quadrature and
basis/derivative evaluation (FP64, numpy).

Restates BESolver/python/basis.py:
  * BSpline.__init__ knot construction, 1.2x extension, threshold snapping
    (basis.py:329-362)
  * BSpline.uniform_knots (basis.py:639-653)
  * BSpline.Gauss_Pn (basis.py:414-429)
  * BSpline.Pn / derivative (basis.py:387-388, 460-497), which go through
    scipy.interpolate.BSpline.basis_element(..., extrapolate=False) followed
    by nan_to_num (zero outside the support).

Pinned by tests/test_oracle_basis.py against scipy.interpolate.BSpline
(basis values and derivatives), partition of unity, and exact polynomial
integration of the quadrature.
"""
import numpy as np


def uniform_knots(k_domain, num_p, sp_order):
    """basis.py:639-653: [k0]*p ++ linspace(k0,k1,num_p-p,endpoint=False) ++ [k1]*(p+1)."""
    t = np.full(sp_order, float(k_domain[0]))
    t = np.append(t, np.linspace(k_domain[0], k_domain[1], num_p - sp_order, endpoint=False))
    t = np.append(t, np.full(sp_order + 1, float(k_domain[1])))
    return t


def log_extended_knots(k_domain, num_p, sp_order, ext_kdomain):
    """basis.py:680-711 (uniform_knots_with_extended_bdy): ceil(num_p/2) uniform knots on
    [k0, k1), then num_p - ceil(num_p/2) - p knots geometric (base 1.2) on [k1, ext), clamped
    at both ends with p and p+1 repeats."""
    num_p1 = int(np.ceil(0.5 * num_p))
    num_p2 = num_p - num_p1
    t = np.full(sp_order, float(k_domain[0]))
    t = np.append(t, np.linspace(k_domain[0], k_domain[1], num_p1, endpoint=False))
    lb = 1.2
    t = np.append(t, np.logspace(np.log(k_domain[1]) / np.log(lb), np.log(ext_kdomain) / np.log(lb),
                                 num_p2 - sp_order, base=lb, endpoint=False))
    return np.append(t, np.full(sp_order + 1, float(ext_kdomain)))


def make_knots(k_domain, sp_order, num_p, sig_pts=None, extend_domain=False, extend_domain_with_log=False):
    """basis.py:329-362 (continuous-Galerkin branch, knots_vec=None).

    Returns (t, domain).  ``extend_domain`` multiplies the upper end by 1.2
    (basis.py:344-346); ``extend_domain_with_log`` keeps half the knots uniform on the
    domain and spaces the rest geometrically up to 5x its end (basis.py:347-349).
    Significant points strictly inside the ORIGINAL k_domain snap the nearest knot
    (argmin) when that knot index lies in (sp_order+1, len(t)-sp_order-1) (basis.py:356-362).
    """
    if extend_domain:
        t = uniform_knots((k_domain[0], 1.2 * k_domain[1]), num_p, sp_order)
        domain = (k_domain[0], 1.2 * k_domain[1])
    elif extend_domain_with_log:
        t = log_extended_knots(k_domain, num_p, sp_order, 5 * k_domain[1])
        domain = (k_domain[0], 5 * k_domain[1])
    else:
        t = uniform_knots(k_domain, num_p, sp_order)
        domain = tuple(k_domain)
    if sig_pts is not None:
        for s in np.asarray(sig_pts, dtype=np.float64).reshape(-1):
            if s > k_domain[0] and s < k_domain[1]:
                idx = int(np.argmin(np.abs(t - s)))
                if (idx > sp_order + 1) and (idx < len(t) - sp_order - 1):
                    t[idx] = s
    return t, domain


def num_knot_intervals(t):
    return len(np.unique(t)) - 1


def gauss_pn(t, q_per_knot):
    """basis.py:414-429: q Gauss-Legendre points on every non-degenerate knot interval."""
    x, w = np.polynomial.legendre.leggauss(q_per_knot)
    qx, qw = [], []
    for i in range(len(t) - 1):
        if t[i] != t[i + 1]:
            qx.append(0.5 * (t[i + 1] - t[i]) * x + 0.5 * (t[i + 1] + t[i]))
            qw.append(0.5 * (t[i + 1] - t[i]) * w)
    qx = np.concatenate(qx)
    qw = np.concatenate(qw)
    assert np.allclose(np.sum(qw), t[-1] - t[0], rtol=1e-14, atol=1e-14)
    return qx, qw


def _basis_element(tk, x):
    """Value of the single B-spline with local knots ``tk`` (len = order+2)
    at points x, zero outside [tk[0], tk[-1]] (scipy basis_element with
    extrapolate=False + nan_to_num).  Cox-de Boor recursion on the local
    knots; the half-open interval convention [t_j, t_{j+1}) is used except
    that x == tk[-1] is assigned to the last non-empty interval, which is
    what scipy's interval search does."""
    x = np.asarray(x, dtype=np.float64)
    deg = len(tk) - 2
    nk = len(tk)
    # order-0 pieces
    N = np.zeros((nk - 1,) + x.shape)
    inside = (x >= tk[0]) & (x <= tk[-1])
    nonempty = [j for j in range(nk - 1) if tk[j] < tk[j + 1]]
    if not nonempty:
        # zero-width element: only ever multiplied by a zero coefficient
        # in basis.py:466-476, so its (nan_to_num'ed) value is irrelevant
        return np.zeros(x.shape)
    last_nonempty = max(nonempty)
    for j in range(nk - 1):
        if tk[j] < tk[j + 1]:
            m = (x >= tk[j]) & (x < tk[j + 1])
            if j == last_nonempty:
                m = m | (x == tk[j + 1])
            N[j] = m
    for d in range(1, deg + 1):
        Nn = np.zeros((nk - 1 - d,) + x.shape)
        for j in range(nk - 1 - d):
            a = tk[j + d] - tk[j]
            b = tk[j + d + 1] - tk[j + 1]
            v = 0.0
            if a != 0:
                v = v + (x - tk[j]) / a * N[j]
            if b != 0:
                v = v + (tk[j + d + 1] - x) / b * N[j + 1]
            Nn[j] = v
        N = Nn
    out = N[0]
    return np.where(inside, out, 0.0)


def basis_eval(t, sp_order, k, x):
    """B_k(x): basis.py:387-388 (Pn) -> basis_element(t[k:k+p+2]) with nan->0."""
    return _basis_element(t[k:k + sp_order + 2], x)


def basis_deriv(t, sp_order, k, x, dorder=1):
    """d^dorder B_k / dx^dorder: basis.py:460-497.

    The reference builds f_alpha(dorder, j) coefficients from knot differences
    (zero where the difference vanishes) and sums lower-order basis elements.
    Only dorder in {1, 2} is used on the hot path."""
    i, p = k, sp_order
    import math
    f1 = math.factorial(p) / math.factorial(p - dorder)

    def f_alpha(kk, j):
        if kk == 0 and j == 0:
            return 1.0
        elif j == 0:
            d1 = t[i + p - kk + 1] - t[i]
            return f_alpha(kk - 1, 0) / d1 if d1 != 0 else 0.0
        elif 0 < j < kk:
            d1 = t[i + p + j - kk + 1] - t[i + j]
            return (f_alpha(kk - 1, j) - f_alpha(kk - 1, j - 1)) / d1 if d1 != 0 else 0.0
        elif kk == j:
            d1 = t[i + p + 1] - t[i + kk]
            return -f_alpha(kk - 1, kk - 1) / d1 if d1 != 0 else 0.0
        raise NotImplementedError

    y = 0.0
    for j in range(dorder + 1):
        tk = t[i + j: i + j + (p - dorder) + 2]
        y = y + f_alpha(dorder, j) * _basis_element(tk, x)
    return y * f1


def vandermonde(t, sp_order, num_p, x, dorder=0):
    """(num_p, len(x)) table of B_k(x) (dorder 0) or B_k'(x) (dorder 1)."""
    if dorder == 0:
        return np.array([basis_eval(t, sp_order, k, x) for k in range(num_p)])
    return np.array([basis_deriv(t, sp_order, k, x, dorder) for k in range(num_p)])
