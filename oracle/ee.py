"""Oracle: electron-electron Fokker-Planck operator in the B-spline x harmonic basis
(FP64, numpy).  TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates (BESolver/python/...):
  * CollisionOpSP.compute_rosenbluth_potentials_op: partial moments
    Pm(m)[i, x] = int_{t_i}^{x} v^m B_i and Qm(m)[i, x] = int_x^{t_(i+4)} v^-m B_i at the
    quadrature points (Gauss orders 4 and 32 x ((p + |m| + 1)//2 + 1) per knot piece),
    the potential kernels m1, m2 per l and hl_op, gl_op = M^-1 [int v^2 B_p m(., k)]
    (collision_operator_spherical.py:786-971), M^-1 applied as a refined solve
    (mass_solve_refined) rather than the reference's explicit inverse;
  * CollisionOpSP.coulomb_collision_op_assembly: cc_a, cc_b[(p,q),(k,l),(r,s)] =
    sum_x w Ia/Ib_{qls}(x) for |k - p|, |r - p| within the reference's window
    (collision_operator_spherical.py:973-1050), with the weak-form terms of the reference's
    own table (cc_terms.py Ia / Ib, read as text into tests/golden/cc_terms_ref.npz by
    tests/golden/make_cc_terms.py; reference_terms() below);
  * the operator cc = M^-1 (cc_a . hl_op + cc_b . gl_op) (bte_0d3v_batched.py:444-476) and
    the Coulomb prefactor gamma_a (bte_0d3v_batched.py:652-661);
  * the e-e residual and Jacobian terms of get_rhs_and_jacobian
    (bte_0d3v_batched.py:702-709, 734-743).

Pinned by tests/test_oracle_ee.py: the term table is the reference's (monomials and values at
seeded points); Pm/Qm against exact integrals; the potentials of a Maxwellian against closed
forms; mass_solve_refined against a long-double residual; the operator conserves
particles, momentum and energy; a Maxwellian is in its null space (C(f_M, f_M) = 0).
"""
import math

import numpy as np
import scipy.constants as sc

from . import bspline
from . import operators as op


def _piece_integral(g, i, m, xb, xe, qx, qw):
    """integrate_xm_splines: sum over the knot pieces of supp(B_i) clipped to [xb, xe]."""
    t = g.t
    p = g.sp_order
    s = 0.0
    for kn in range(i + 1, i + p + 2):
        a = max(xb, t[kn - 1])
        b = min(xe, t[kn])
        if b > a:
            x = 0.5 * ((b - a) * qx + (b + a))
            w = 0.5 * (b - a) * qw
            s += np.dot(w, x ** m * bspline.basis_eval(t, p, i, x))
    return s


def Pm(g, m, gx):
    p = g.sp_order
    q = ((p + abs(m) + 1) // 2 + 1) * 4
    qx, qw = np.polynomial.legendre.leggauss(q)
    t = g.t
    out = np.zeros((g.num_p, len(gx)))
    for i in range(g.num_p):
        lo, hi = t[i], t[i + p + 1]
        for j in np.nonzero((gx >= lo) & (gx <= hi))[0]:
            out[i, j] = _piece_integral(g, i, m, lo, gx[j], qx, qw)
        out[i, gx > hi] = _piece_integral(g, i, m, lo, hi, qx, qw)
    return out


def Qm(g, m, gx):
    p = g.sp_order
    q = ((p + abs(m) + 1) // 2 + 1) * 32
    qx, qw = np.polynomial.legendre.leggauss(q)
    t = g.t
    out = np.zeros((g.num_p, len(gx)))
    for i in range(g.num_p):
        lo, hi = t[i], t[i + p + 1]
        for j in np.nonzero((gx >= lo) & (gx <= hi))[0]:
            out[i, j] = _piece_integral(g, i, -m, gx[j], hi, qx, qw)
        out[i, gx < lo] = _piece_integral(g, i, -m, lo, hi, qx, qw)
    return out


def mass_solve_refined(Mr, raw, steps=3):
    """Mr^-1 raw for the radial mass block: a Cholesky solve, then iterative refinement on
    the residual formed in long double.  The result is the solution for the FP64 Mr and raw
    to about one rounding; the reference's product with the explicit inverse
    (collision_operator_spherical.py:966-967) carries cond(Mr) roundings, ~1e-14 of the
    largest entry at Nr = 128, and that moves e-e dominated tail rates by ~2e-8."""
    import scipy.linalg as sla
    cf = sla.cho_factor(Mr, lower=True)
    y = sla.cho_solve(cf, raw)
    LD = np.longdouble
    ML, RL = Mr.astype(LD), raw.astype(LD)
    for _ in range(steps):
        r = (RL - ML @ y.astype(LD)).astype(np.float64)
        y = y + sla.cho_solve(cf, r)
    return y


def rosenbluth_ops(g: op.Grid, Minv=None, m_ab=1.0):
    """(hl_op, gl_op), each (n, n), block diagonal in l, premultiplied by M^-1 (applied per
    l block as a refined solve with the radial mass block; Minv is not used)."""
    gx, gw = g.quad(g.q)
    V = g.V(gx)
    ns = g.num_sh
    Mr = op.mass_matrix(g)[0::ns, 0::ns]
    hl = np.zeros((g.n, g.n))
    gl = np.zeros((g.n, g.n))
    for l in range(ns):
        f1 = (l - 0.5) / (l + 1.5)
        m1 = (2.0 / (2 * l + 1)) * (1 + m_ab) * (Pm(g, l + 2, gx) / gx ** (l + 1) + Qm(g, l - 1, gx) * gx ** l)
        m2 = -(2.0 / (4 * l * l - 1)) * ((Pm(g, l + 2, gx) / gx ** (l - 1) - f1 * Pm(g, l + 4, gx) / gx ** (l + 1))
                                         + (Qm(g, l - 3, gx) * gx ** l - f1 * Qm(g, l - 1, gx) * gx ** (l + 2)))
        hl[l::ns, l::ns] = mass_solve_refined(Mr, (V * (gw * gx ** 2)) @ m1.T)
        gl[l::ns, l::ns] = mass_solve_refined(Mr, (V * (gw * gx ** 2)) @ m2.T)
    return hl, gl


_REF_TERMS = None


def reference_terms():
    """The reference's weak-form term table (cc_terms.py Ia / Ib for (q, l, s) <= 4) from the
    committed golden data: {'a'|'b': {(q, l, s): [(coef, pow v, dp, dk, dr), ...]}}."""
    global _REF_TERMS
    if _REF_TERMS is None:
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "tests", "golden", "cc_terms_ref.npz")
        d = np.load(path)
        out = {}
        for kind in ("a", "b"):
            tab = {}
            for row in d[f"terms_{kind}"]:
                key = (int(row[0]), int(row[1]), int(row[2]))
                tab.setdefault(key, []).append((float(row[7]), int(row[3]), int(row[4]), int(row[5]), int(row[6])))
            out[kind] = tab
        _REF_TERMS = out
    return _REF_TERMS


def cc_tensors(g: op.Grid, terms=None, lmax=None):
    """Dense cc_a, cc_b (n, n, n) (small grids only); terms: the reference's table by default."""
    terms = reference_terms() if terms is None else terms
    ns = g.num_sh
    lmax = g.lmax if lmax is None else lmax
    p_ = g.sp_order
    t = g.t
    gx_a, gw_a = g.quad(g.q)
    n = g.n
    A = np.zeros((n, n, n))
    Bt = np.zeros((n, n, n))
    D = [g.V(gx_a, 0), g.V(gx_a, 1), bspline_second(g, gx_a)]
    xL = gx_a.astype(np.longdouble)
    win = p_ + 3
    for p in range(g.num_p):
        for k in range(max(0, p - win), min(g.num_p, p + win)):
            for r in range(max(0, p - win), min(g.num_p, p + win)):
                kmin = min(t[p], t[k], t[r])
                kmax = max(t[r + p_ + 1], t[p + p_ + 1], t[k + p_ + 1])
                sel = (gx_a >= kmin) & (gx_a <= kmax)
                x, w = gx_a[sel], gw_a[sel]
                for key, T in (("a", A), ("b", Bt)):
                    for (q, l, s), mons in terms[key].items():
                        if q > lmax or l > lmax or s > lmax:
                            continue
                        # long-double sums: the monomials of one combination have both signs
                        val = np.zeros(x.shape, dtype=np.longdouble)
                        for c, pw, dp, dk, dr in mons:
                            val = val + np.longdouble(c) * xL[sel] ** pw * (D[dp][p][sel] * D[dk][k][sel]).astype(
                                np.longdouble) * D[dr][r][sel]
                        T[p * ns + q, k * ns + l, r * ns + s] = float(np.dot(w.astype(np.longdouble), val))
    return A, Bt


def bspline_second(g, x):
    return np.array([bspline.basis_deriv(g.t, g.sp_order, k, x, 2) for k in range(g.num_p)])


def ee_operator(Minv, cc_a, cc_b, hl, gl):
    """cc = M^-1 (cc_a . hl_op + cc_b . gl_op) (tensordot over the last index)."""
    n = Minv.shape[0]
    cc = np.tensordot(cc_a, hl, axes=([2], [0])) + np.tensordot(cc_b, gl, axes=([2], [0]))
    return (Minv @ cc.reshape(n, -1)).reshape(n, n, n)


def gamma_a(ops, x, ne):
    """bte_0d3v_batched.py:652-661: Coulomb prefactor from the current distribution."""
    vth = ops.g.vth
    mw0 = op.maxwellian0(vth)
    Mop = ops.mass
    Top = ops.temp * op.TEMP_K_1EV
    m0 = mw0 * (Mop @ x) * vth ** 3
    kT = np.abs(mw0 * (Top @ x) / m0 * vth ** 5 * sc.Boltzmann).reshape(-1)
    eps0, qe, me = sc.epsilon_0, sc.e, sc.electron_mass
    c_lambda = (12 * np.pi * (eps0 * kT) ** 1.5) / (qe ** 3 * np.sqrt(ne))
    return np.log(c_lambda) * qe ** 4 / (4 * np.pi * (eps0 * me) ** 2) / vth ** 3


def ee_term(cc, x):
    """(cc . x) . x per column of x: c_ee_xx of bte_0d3v_batched.py:702-703."""
    ccx = np.tensordot(cc, x, axes=([2], [0]))           # (n, n, B)
    return np.einsum("abc,bc->ac", ccx, x)
