"""Writes tests/golden/bspline_exact.npz: exact (rational, sympy) B-spline integrals.

The cubic B-splines of two small knot vectors on [0, 3] (a uniform clamped one,
and the same with one interior knot moved, as threshold snapping does) are
built with sympy.bspline_basis_set (Cox-de Boor in exact arithmetic), and the
radial integrals every Galerkin operator of the velocity-space path is made of
are integrated piece by piece in rationals:

  mass    M[p,k]  = int x^2 B_p B_k          mm1[p,k] = int x^2 B_p B_k'
  mm2     [p,k]   = int x   B_p B_k          S3[p,k]  = int x^3 B_p B_k
  D3      [p,k]   = int x^3 B_p' B_k'        m2/m3/m4[k] = int x^{2,3,4} B_k
plus values B_k, B_k', B_k'' at a few points.  These are the data the oracle is
checked against; nothing here depends on the oracle.

Run:  python tests/golden/make_golden.py      (a few seconds)
"""
import os

import numpy as np
import sympy as sp

HERE = os.path.dirname(os.path.abspath(__file__))
X = sp.Symbol("x", real=True)


def knot_vectors():
    h = sp.Rational(3, 5)
    uni = [0, 0, 0] + [i * h for i in range(5)] + [3, 3, 3, 3]
    snapped = list(uni)
    snapped[5] = sp.Rational(11, 10)      # interior knot 1.2 -> 1.1
    return {"uniform": [sp.nsimplify(t) for t in uni], "snapped": [sp.nsimplify(t) for t in snapped]}


def pieces(knots):
    """pieces[k][i] = polynomial of B_k on (u_i, u_{i+1}) of the distinct knots u."""
    basis = sp.bspline_basis_set(3, knots, X)
    u = sorted(set(knots))
    out = []
    for Bk in basis:
        row = []
        for a, b in zip(u[:-1], u[1:]):
            mid = (a + b) / 2
            if isinstance(Bk, sp.Piecewise):
                expr = 0
                for e, c in Bk.args:
                    if bool(c.subs(X, mid)):
                        expr = e
                        break
            else:
                expr = Bk
            row.append(sp.expand(expr))
        out.append(row)
    return u, out


def integrals(knots):
    u, P = pieces(knots)
    npk = len(P)
    dP = [[sp.diff(e, X) for e in row] for row in P]
    res = {k: np.zeros((npk, npk)) for k in ("M", "mm1", "mm2", "S3", "D3")}
    for k in ("m2", "m3", "m4"):
        res[k] = np.zeros(npk)
    for i, (a, b) in enumerate(zip(u[:-1], u[1:])):
        for p in range(npk):
            bp, dbp = P[p][i], dP[p][i]
            if bp == 0:
                continue
            for pw, key in ((2, "m2"), (3, "m3"), (4, "m4")):
                res[key][p] += float(sp.integrate(X ** pw * bp, (X, a, b)))
            for k in range(npk):
                bk, dbk = P[k][i], dP[k][i]
                if bk == 0:
                    continue
                res["M"][p, k] += float(sp.integrate(X ** 2 * bp * bk, (X, a, b)))
                res["mm1"][p, k] += float(sp.integrate(X ** 2 * bp * dbk, (X, a, b)))
                res["mm2"][p, k] += float(sp.integrate(X * bp * bk, (X, a, b)))
                res["S3"][p, k] += float(sp.integrate(X ** 3 * bp * bk, (X, a, b)))
                res["D3"][p, k] += float(sp.integrate(X ** 3 * dbp * dbk, (X, a, b)))
    # point values (left-closed pieces; the last point is the right end of the domain)
    pts = [sp.Rational(1, 7), sp.Rational(3, 5), sp.Rational(13, 10), sp.Rational(29, 11), sp.Integer(3)]
    vals = np.zeros((3, npk, len(pts)))
    for j, xp in enumerate(pts):
        idx = max(i for i in range(len(u) - 1) if u[i] <= xp) if xp < u[-1] else len(u) - 2
        for p in range(npk):
            e = P[p][idx]
            vals[0, p, j] = float(e.subs(X, xp))
            vals[1, p, j] = float(sp.diff(e, X).subs(X, xp))
            vals[2, p, j] = float(sp.diff(e, X, 2).subs(X, xp))
    res["pts"] = np.array([float(v) for v in pts])
    res["vals"] = vals
    res["knots"] = np.array([float(t) for t in knots])
    return res


def main():
    out = {}
    for name, kv in knot_vectors().items():
        for key, val in integrals(kv).items():
            out[f"{name}_{key}"] = val
    path = os.path.join(HERE, "bspline_exact.npz")
    np.savez(path, **out)
    print("wrote", path, sorted(out))


if __name__ == "__main__":
    main()
