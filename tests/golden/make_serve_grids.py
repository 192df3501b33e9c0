"""Writes tests/golden/serve_grids.npz: oracle QoIs of serving requests on every grid of the
serving mix (Nr in {64, 128, 256} x lmax in {1, 2, 4}, three processes), two requests per grid
without e-e, and one request with e-e (ionisation degree 1e-3) on each grid where the dense
e-e oracle is affordable here (n <= 325: Nr 64 at lmax 1, 2, 4 and Nr 128 at lmax 1).  The
e-e request of the Nr = 128, lmax = 2 grid is covered by ee_n387.npz.

Without e-e: the oracle's exact Newton polished by refine_fixed_point.  With e-e: the oracle's
dense exact Newton (e-e Jacobian cc x + cc_12 x) from the converged no-e-e state to 1e-15.

Run from the repository root:  python tests/golden/make_serve_grids.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from helpers import oracle_batched, oracle_fixed_point  # noqa: E402
from oracle import ee as eeo, solver0d as s0  # noqa: E402

GRIDS = [(Nr, lmax) for Nr in (64, 128, 256) for lmax in (1, 2, 4)]
EE_GRIDS = [(64, 1), (64, 2), (64, 4), (128, 1)]
N0, TG = 3.22e22, 300.0


def requests():
    rng = np.random.default_rng(20251203)
    out = []
    for Nr, lmax in GRIDS:
        for _ in range(2):
            out.append((Nr, lmax, float(10.0 ** rng.uniform(0.0, 3.0)), 0.0, rng.uniform(0.8, 1.2, 3)))
    for Nr, lmax in EE_GRIDS:
        out.append((Nr, lmax, float(10.0 ** rng.uniform(0.5, 2.5)), 1e-3, rng.uniform(0.8, 1.2, 3)))
    return out


def main():
    rows = []
    for Nr, lmax, EbyN, ion, fac in requests():
        ops, g, procs = oracle_batched(Nr, lmax)
        prm = {"n0": np.array([N0]), "Tg": np.array([TG]), "E": np.array([EbyN * N0 * 1e-21]),
               "ns_by_n0": np.asarray(fac, dtype=np.float64).reshape(3, 1)}
        h, q, info = oracle_fixed_point(ops, prm)
        rel = float(info[0]["rel"][0])
        if ion > 0:
            prm["ne"] = prm["n0"] * ion
            hl, gl = eeo.rosenbluth_ops(g)
            A, Bt = eeo.cc_tensors(g)
            ops.ee = {"cc": eeo.ee_operator(ops.Minv, A, Bt, hl, gl)}
            del A, Bt
            _, q, info = s0.steady_state_solve_each(ops, prm, h, 0.0, 1e-15, 60, jacobian="exact")
            rel = float(info[0]["rel"][0])
        print(f"Nr {Nr} lmax {lmax} E/N {EbyN:.4g} ion {ion:g}: rel {rel:.2e} energy {float(q['energy'][0]):.10g}",
              flush=True)
        rows.append((Nr, lmax, EbyN, ion, *fac, float(q["energy"][0]), float(q["mobility"][0]),
                     float(q["diffusion"][0]), *[float(v) for v in q["rates"][:, 0]], rel))
    np.savez(os.path.join(HERE, "serve_grids.npz"), rows=np.array(rows),
             columns=np.array(["Nr", "lmax", "EbyN", "ion_deg", "fac0", "fac1", "fac2", "energy", "mobility",
                               "diffusion", "rate0", "rate1", "rate2", "rel"]))


if __name__ == "__main__":
    main()
