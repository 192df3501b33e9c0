"""Golden data for the e-e weak-form term table: the reference's Ia / Ib integrands
(BESolver/python/cc_terms.py, generated there with sympy) read as TEXT and turned into data.

The file is parsed with `ast` (nothing of it is imported, executed or evaluated by Python):
every `if (q,l,s) == (a,b,c): return <expr>` branch of Ia and Ib is converted node by node
into a sympy expression in the symbols of the integrand (vr, B_*_vr, DB_*_dvr,
DB_*_dvr_dvr), expanded into monomials  c * vr^a * B_p^(dp) B_k^(dk) B_r^(dr),  and written
as a table, together with the values of every branch at seeded random points.  Only
(q, l, s) with all indices <= LMAX are kept.

    python tests/golden/make_cc_terms.py /root/reference/BESolver/python/cc_terms.py

writes tests/golden/cc_terms_ref.npz (run in the build container; the reference is not
needed afterwards).
"""
import ast
import os
import sys

import numpy as np
import sympy as sp

LMAX = 4
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "cc_terms_ref.npz")

VR = sp.Symbol("vr", positive=True)
# radial factor symbols: name -> (function index 0 p / 1 k / 2 r, derivative order)
FACTORS = {}
for a, ia in (("p", 0), ("k", 1), ("r", 2)):
    FACTORS[f"B_{a}_vr"] = (ia, 0)
    FACTORS[f"DB_{a}_dvr"] = (ia, 1)
    FACTORS[f"DB_{a}_dvr_dvr"] = (ia, 2)
SYMS = {name: sp.Symbol(name, real=True) for name in FACTORS}
SYMS["vr"] = VR


def to_sympy(node):
    """AST expression node -> sympy expression (numbers, the integrand symbols, + - * / **,
    unary minus, math.sqrt(.) and math.pi only)."""
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
        return sp.nsimplify(node.value) if isinstance(node.value, float) else sp.Integer(node.value)
    if isinstance(node, ast.Name):
        if node.id not in SYMS:
            raise ValueError(f"unexpected name {node.id}")
        return SYMS[node.id]
    if isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name) and node.value.id == "math":
        if node.attr == "pi":
            return sp.pi
        raise ValueError(f"unexpected attribute math.{node.attr}")
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        return -to_sympy(node.operand)
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.UAdd):
        return to_sympy(node.operand)
    if isinstance(node, ast.BinOp):
        a, b = to_sympy(node.left), to_sympy(node.right)
        if isinstance(node.op, ast.Add):
            return a + b
        if isinstance(node.op, ast.Sub):
            return a - b
        if isinstance(node.op, ast.Mult):
            return a * b
        if isinstance(node.op, ast.Div):
            return a / b
        if isinstance(node.op, ast.Pow):
            return a ** b
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Attribute) and node.func.attr == "sqrt" \
            and isinstance(node.func.value, ast.Name) and node.func.value.id == "math" and len(node.args) == 1:
        return sp.sqrt(to_sympy(node.args[0]))
    raise ValueError(f"unsupported expression node {ast.dump(node)[:120]}")


def branches(tree, fname):
    """{(q, l, s): sympy expression} of the if-branches of function fname."""
    fn = next(n for n in tree.body if isinstance(n, ast.FunctionDef) and n.name == fname)
    out = {}
    for st in fn.body:
        if not isinstance(st, ast.If):
            continue
        cmp = st.test
        if not (isinstance(cmp, ast.Compare) and isinstance(cmp.comparators[0], ast.Tuple)):
            continue
        key = tuple(int(e.value) for e in cmp.comparators[0].elts)
        ret = next(s for s in st.body if isinstance(s, ast.Return))
        out[key] = to_sympy(ret.value)
    return out


def monomials(expr):
    """Expanded expr as a list of (coefficient, power of vr, dp, dk, dr)."""
    out = []
    for term in sp.Add.make_args(sp.expand(expr)):
        coef = sp.Integer(1)
        pw = 0
        der = [None, None, None]
        for f in sp.Mul.make_args(term):
            base, ex = f.as_base_exp()
            if base == VR:
                pw += int(ex)
            elif isinstance(base, sp.Symbol) and base.name in FACTORS:
                ia, d = FACTORS[base.name]
                if ex != 1 or der[ia] is not None:
                    raise ValueError(f"factor {f} in {term}")
                der[ia] = d
            else:
                coef = coef * f
        if any(d is None for d in der):
            raise ValueError(f"term {term} lacks a radial factor")
        out.append((float(coef), pw, der[0], der[1], der[2]))
    return out


def main(path):
    tree = ast.parse(open(path).read())
    rng = np.random.default_rng(20251203)
    npts = 16
    pts = {name: rng.uniform(0.5, 2.0, npts) for name in SYMS}
    data = {"points_names": np.array(sorted(pts)), "points": np.array([pts[k] for k in sorted(pts)])}
    for kind, fname in (("a", "Ia"), ("b", "Ib")):
        br = branches(tree, fname)
        keys = sorted(k for k in br if max(k) <= LMAX)
        rows, vals = [], []
        for k in keys:
            mons = monomials(br[k])
            for m in mons:
                rows.append((k[0], k[1], k[2]) + m[1:] + (m[0],))
            f = sp.lambdify([SYMS[n] for n in sorted(pts)], br[k], "numpy")
            vals.append(np.broadcast_to(np.asarray(f(*[pts[n] for n in sorted(pts)]), dtype=np.float64), (npts,)))
        rows = np.array(rows, dtype=np.float64)
        data[f"terms_{kind}"] = rows              # q l s pow dp dk dr coef
        data[f"keys_{kind}"] = np.array(keys, dtype=np.int64)
        data[f"values_{kind}"] = np.array(vals)
        print(f"{fname}: {len(keys)} (q,l,s) with indices <= {LMAX}, {len(rows)} monomials")
    np.savez(OUT, **data)
    print("wrote", OUT)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/BESolver/python/cc_terms.py")
