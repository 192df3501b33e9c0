"""Pins of the oracle's B-spline layer (knots, basis, derivatives, quadrature).

References restated by oracle/bspline.py: basis.py:329-362, 387-388, 414-429,
460-497, 639-653.  Checked here against
  * exact rational values of sympy's Cox-de Boor basis (tests/golden/bspline_exact.npz,
    written by tests/golden/make_golden.py), uniform and snapped knots;
  * scipy.interpolate.BSpline (values, first and second derivatives) on the knot
    vectors the solvers actually use (1.2x extension and threshold snapping);
  * partition of unity sum_p B_p = 1 and sum_p B_p' = 0;
  * exactness of q-point Gauss-Legendre per interval for degree 2q-1 and not 2q.
"""
import os

import numpy as np
import pytest
import scipy.interpolate

from oracle import bspline

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "bspline_exact.npz"))


@pytest.mark.parametrize("name", ["uniform", "snapped"])
def test_basis_matches_exact_rational_values(name):
    t = GOLD[f"{name}_knots"]
    pts = GOLD[f"{name}_pts"]
    vals = GOLD[f"{name}_vals"]
    npk = len(t) - 4
    for k in range(npk):
        np.testing.assert_allclose(bspline.basis_eval(t, 3, k, pts), vals[0, k], rtol=0, atol=1e-14)
        np.testing.assert_allclose(bspline.basis_deriv(t, 3, k, pts, 1), vals[1, k], rtol=0, atol=1e-13)
        np.testing.assert_allclose(bspline.basis_deriv(t, 3, k, pts, 2), vals[2, k], rtol=0, atol=1e-12)


def solver_knots():
    kd = (0.0, 6.7)
    t_plain = bspline.uniform_knots(kd, 33, 3)
    t_snap, _ = bspline.make_knots(kd, 3, 33, sig_pts=np.array([2.53, 4.71]), extend_domain=True)
    return [t_plain, t_snap]


@pytest.mark.parametrize("which", [0, 1])
def test_basis_matches_scipy_bspline(which):
    t = solver_knots()[which]
    npk = len(t) - 4
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(t[0], t[-1], 400), np.unique(t)])
    for k in range(npk):
        c = np.zeros(npk)
        c[k] = 1.0
        spl = scipy.interpolate.BSpline(t, c, 3, extrapolate=False)
        for nu in (0, 1, 2):
            ref = np.nan_to_num(spl(x, nu=nu))
            got = bspline.basis_eval(t, 3, k, x) if nu == 0 else bspline.basis_deriv(t, 3, k, x, nu)
            scale = max(1.0, np.abs(ref).max())
            # inside the support the two agree to rounding; at interior knots the
            # second derivative jumps and both pick the right-hand piece
            np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * scale)


def test_partition_of_unity():
    for t in solver_knots():
        x = np.linspace(t[0], t[-1], 997)
        V = bspline.vandermonde(t, 3, len(t) - 4, x)
        D = bspline.vandermonde(t, 3, len(t) - 4, x, 1)
        np.testing.assert_allclose(V.sum(axis=0), 1.0, rtol=0, atol=1e-14)
        np.testing.assert_allclose(D.sum(axis=0), 0.0, rtol=0, atol=1e-12 * np.abs(D).max())


def test_knot_vector_structure():
    kd = (0.0, 6.7)
    t = bspline.uniform_knots(kd, 33, 3)
    assert len(t) == 33 + 4
    assert (t[:4] == 0.0).all() and (t[-4:] == 6.7).all()
    inner = np.diff(t[3:-3])
    np.testing.assert_allclose(inner, 6.7 / 30, rtol=1e-13)
    te, dom = bspline.make_knots(kd, 3, 33, extend_domain=True)
    assert dom[1] == pytest.approx(1.2 * 6.7) and te[-1] == pytest.approx(1.2 * 6.7)
    # snapping moves exactly the nearest interior knot onto each significant point;
    # points outside the un-extended domain are ignored
    ts, _ = bspline.make_knots(kd, 3, 33, sig_pts=np.array([2.53, 7.9]), extend_domain=True)
    moved = np.nonzero(ts != te)[0]
    assert len(moved) == 1 and ts[moved[0]] == 2.53
    assert moved[0] == np.argmin(np.abs(te - 2.53))


@pytest.mark.parametrize("num_p", [33, 64, 129])
def test_log_extended_knot_vector(num_p):
    """LOG extension (basis.py:347-349, 680-711): p + 1 clamped knots at each end, ceil(num_p/2)
    knots with constant spacing k1 / ceil(num_p/2) from k0 = 0, then a geometric run from k1
    whose ratio is 5^(1 / (num_p - ceil(num_p/2) - p)) up to the clamped end 5 k1; the knot
    count fixes num_p basis functions; snapping works inside the un-extended domain."""
    kd, p = (0.0, 6.7), 3
    t, dom = bspline.make_knots(kd, p, num_p, extend_domain_with_log=True)
    n1 = -(-num_p // 2)
    n2 = num_p - n1
    assert len(t) == num_p + p + 1 and dom == (0.0, 5 * 6.7)
    assert (t[:p + 1] == 0.0).all() and np.allclose(t[-(p + 1):], 5 * 6.7, rtol=1e-14, atol=0)
    np.testing.assert_allclose(np.diff(t[p:p + n1 + 1]), 6.7 / n1, rtol=1e-12)
    geo = t[p + n1:p + n1 + (n2 - p) + 1]          # k1, ..., 5 k1
    assert geo[0] == pytest.approx(6.7, rel=1e-14)
    np.testing.assert_allclose(geo[1:] / geo[:-1], 5.0 ** (1.0 / (n2 - p)), rtol=1e-12)
    ts, _ = bspline.make_knots(kd, p, num_p, sig_pts=np.array([2.53, 9.0]), extend_domain_with_log=True)
    moved = np.nonzero(ts != t)[0]
    assert len(moved) == 1 and ts[moved[0]] == 2.53


@pytest.mark.parametrize("q", [3, 5, 10])
def test_gauss_points_per_interval(q):
    for t in solver_knots():
        x, w = bspline.gauss_pn(t, q)
        assert len(x) == q * (len(np.unique(t)) - 1)
        a, b = t[0], t[-1]
        for m in range(0, 2 * q):
            exact = (b ** (m + 1) - a ** (m + 1)) / (m + 1)
            assert abs(np.dot(w, x ** m) - exact) <= 1e-12 * abs(exact)
    # degree 2q is no longer integrated exactly (one unit interval, where the
    # Gauss error (q!)^4 / ((2q+1) ((2q)!)^2) is far above rounding)
    x, w = bspline.gauss_pn(np.array([0.0] * 4 + [1.0] * 4), q)
    assert abs(np.dot(w, x ** (2 * q)) - 1.0 / (2 * q + 1)) > 1e-13
