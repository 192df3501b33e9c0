"""Writes tests/golden/ee_n387.npz: the oracle's e-e steady state (exact Newton with the dense
e-e Jacobian to the FP64 floor) of four seeded instances on the Nr = 128, lmax = 2 grid
(n = 387), the size of the batched e-e workload.  The dense oracle takes minutes per instance
at this size, so its QoIs are stored here and the GPU test compares against them.

Run from the repository root:  python tests/golden/make_ee387.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from helpers import instance_params, oracle_batched  # noqa: E402
from oracle import ee as eeo, solver0d as s0  # noqa: E402


def main():
    ops, g, procs = oracle_batched(128, 2)
    hl, gl = eeo.rosenbluth_ops(g, ops.Minv)
    A, Bt = eeo.cc_tensors(g)
    cc = eeo.ee_operator(ops.Minv, A, Bt, hl, gl)
    B = 4
    prm = instance_params(B, seed=387)
    prm["ne"] = prm["n0"] * np.array([1e-2, 1e-3, 1e-4, 1e-6])
    x0, _, _ = s0.steady_state_solve_each(ops, prm, ops.initial(B), 0.0, 1e-14, 40, jacobian="exact")
    ops.ee = {"cc": cc}
    x, q, info = s0.steady_state_solve_each(ops, prm, x0, 0.0, 1e-15, 40, jacobian="exact")
    print("iterations", [i["iters"] for i in info], "rel", [float(i["rel"][0]) for i in info])
    np.savez(os.path.join(HERE, "ee_n387.npz"), n0=prm["n0"], Tg=prm["Tg"], E=prm["E"], ns_by_n0=prm["ns_by_n0"],
             ne=prm["ne"], energy=q["energy"], mobility=q["mobility"], diffusion=q["diffusion"], rates=q["rates"],
             rel=np.array([float(i["rel"][0]) for i in info]))


if __name__ == "__main__":
    main()
