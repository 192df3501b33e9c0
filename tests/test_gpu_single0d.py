"""GPU parity of the single-instance drop-in (bte_0d3v) against the oracle's single path.

Oracle: bte_0d3v_solver.steady_state_solver + compute_QoIs restated in
oracle/solver0d.py (newton_solve with the reference Jacobian and step halving,
rtol 1e-13 on the relative change, atol 1e-8).  Both solvers stop near the FP64
floor, so the QoIs agree to 1e-9 relative (rates with the 1e-9 floor rule).
"""
import argparse

import numpy as np
import pytest

from conftest import XSEC_FILE
from helpers import oracle_procs, qoi_errors, rel_err
from oracle import operators as op
from oracle import solver0d as s0

pytestmark = pytest.mark.gpu


def _args(E_by_N, Nr=40, lmax=1):
    return argparse.Namespace(collisions=XSEC_FILE, n0=3.22e22, Tg=300.0, E_field=E_by_N * 3.22e22 * 1e-21,
                              electron_volt=4.0, ev_max=80.0, Nr=Nr, l_max=lmax, sp_order=3, spline_qpts=5,
                              ns_by_n0=[1.0, 1.0, 1.0], use_dg=0, vstype=0, ee_collisions=0, efield_period=0.0, ion_deg=0.0)


def _oracle(args):
    g = op.grid_single(args.Nr, args.l_max, args.electron_volt, args.ev_max, args.spline_qpts)
    ops = s0.SingleOps(g, oracle_procs(), args.n0, args.Tg)
    r = s0.single_steady_state(ops, args.E_field)
    return r, s0.single_qois(ops, r["sol"][1], args.E_field)


@pytest.mark.parametrize("E_by_N,jac", [(30.0, "reference"), (300.0, "reference"), (100.0, "exact")])
def test_single_steady_state_matches_oracle(E_by_N, jac):
    from boltzmann_amd.bte_0d3v import bte_0d3v
    print(f"case: single 0D E/N={E_by_N} Td jacobian={jac}", flush=True)
    args = _args(E_by_N)
    solver = bte_0d3v(args)
    solver.setup()
    out = solver.steady_state_solver(jacobian=jac)
    q = solver.compute_QoIs(out["sol"][1], None)
    r_ref, q_ref = _oracle(args)
    assert out["sol"].shape == r_ref["sol"].shape
    assert rel_err(out["sol"][0], r_ref["sol"][0]) < 1e-13         # normalised initial data
    assert rel_err(out["sol"][1], r_ref["sol"][1]) < 1e-9
    qa = {k: np.asarray(q[k]).reshape(-1) for k in ("energy", "mobility", "diffusion")}
    qa["rates"] = np.asarray(q["rates"]).reshape(len(q["rates"]), -1)
    qb = {k: np.asarray(q_ref[k]).reshape(-1) for k in ("energy", "mobility", "diffusion")}
    qb["rates"] = np.asarray(q_ref["rates"]).reshape(len(q_ref["rates"]), -1)
    errs = qoi_errors(qa, qb)
    assert max(errs.values()) < 1e-9, errs


def test_config1_exact_against_eigenpair():
    """The CPU configuration exactly: Nr = 64, lmax = 1, E/N = 100 Td, Tg = 300 K, q = 5,
    against the eigenpair oracle.  The single-path steady state is the eigenvector of the
    dominant (largest real part) generalized eigenvalue of (C + E A, M), normalised to
    u.h = 1 (scipy.linalg.eig on the oracle's operators); the GPU solve must match that
    vector and its QoIs to 1e-8."""
    import scipy.linalg
    from boltzmann_amd.bte_0d3v import bte_0d3v
    print("case: single 0D config 1 (Nr=64 lmax=1 100 Td) vs eigenpair", flush=True)
    args = _args(100.0, Nr=64, lmax=1)
    solver = bte_0d3v(args)
    solver.setup()
    out = solver.steady_state_solver(jacobian="exact")
    h = np.asarray(out["sol"][1], dtype=np.float64)
    g = op.grid_single(64, 1, args.electron_volt, args.ev_max, args.spline_qpts)
    ops = s0.SingleOps(g, oracle_procs(), args.n0, args.Tg)
    K = ops.C + ops.Araw * (args.E_field / g.vth) * (op.QE / op.ME)
    ev, V = scipy.linalg.eig(K, ops.M)
    # the eigenvalue is the growth rate of the oracle's own single-path fixed point (the
    # dominant mode; spurious modes of the truncated domain can have larger real parts
    # within the rounding of the largest |eigenvalue|)
    r = s0.single_steady_state(ops, args.E_field)
    hr = r["sol"][1] / (ops.u @ r["sol"][1])
    lam = (ops.u @ (ops.Minv @ (K @ hr)))
    fin = np.isfinite(ev)
    k = np.nonzero(fin)[0][np.argmin(np.abs(ev[fin] - lam))]
    v = V[:, k]
    v = (v / v[np.argmax(np.abs(v))]).real
    v = v / (ops.u @ v)
    assert abs(ev[k].imag) <= 1e-10 * abs(ev[k])
    hn = h / (ops.u @ h)
    assert rel_err(hn, v) < 1e-8
    q = solver.compute_QoIs(out["sol"][1], None)
    q_ref = s0.single_qois(ops, v * (ops.u @ h), args.E_field)       # scaled as the GPU solution
    qa = {k_: np.asarray(q[k_]).reshape(-1) for k_ in ("energy", "mobility", "diffusion")}
    qa["rates"] = np.asarray(q["rates"]).reshape(len(q["rates"]), -1)
    qb = {k_: np.asarray(q_ref[k_]).reshape(-1) for k_ in ("energy", "mobility", "diffusion")}
    qb["rates"] = np.asarray(q_ref["rates"]).reshape(len(q_ref["rates"]), -1)
    errs = qoi_errors(qa, qb)
    print("config 1 qoi errors vs eigenpair", errs, flush=True)
    assert max(errs.values()) < 1e-8, errs
