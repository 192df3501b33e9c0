"""Pins of the oracle's 0D solvers (single and batched) and QoI assembly.

References restated by oracle/solver0d.py: bte_0d3v_solver.py:24-73, 404-460,
741-859, 1234-1276 and bte_0d3v_batched.py:425-442, 602-821, 882-979, 1163-1217.
Checked here against what the mathematics and physics fix:
  * Q is an orthonormal basis of u-perp (any such basis gives the same h);
  * the single-path fixed point is the dominant generalized eigenpair
    (C + E A) h = lambda M h, u^T h = 1 (scipy.linalg.eig);
  * the batched fixed point satisfies K x - mu M x = c M uhat with mu = w . x
    (the batched growth vector omits advection), for both Jacobians;
  * elastic collisions with the gas-temperature term at vanishing field relax to
    the Maxwellian at Tg: mean energy 3/2 k Tg, and the mobility / diffusion ratio
    obeys the Einstein relation mu / D = e / (k Tg).  This fixes the elastic
    drag and diffusion coefficients, the sign and size of the field coupling and
    the normalisations of the mobility and diffusion QoIs.
"""
import math

import numpy as np
import pytest
import scipy.linalg

from helpers import instance_params, oracle_batched, oracle_procs
from oracle import operators as op
from oracle import solver0d as s0


def test_qr_basis_spans_u_perp():
    ops, g, _ = oracle_batched(24, 2)
    Q, u = ops.Q, ops.u
    np.testing.assert_allclose(Q.T @ Q, np.eye(g.n - 1), atol=1e-12)
    assert np.abs(Q.T @ u).max() < 1e-12 * np.linalg.norm(u)


@pytest.fixture(scope="module")
def single_case():
    procs = oracle_procs()
    g = op.grid_single(40, 1, 4.0, 80.0, 5)
    ops = s0.SingleOps(g, procs, 3.22e22, 300.0)
    E = 100.0 * 1e-21 * 3.22e22
    r = s0.single_steady_state(ops, E)
    return ops, g, E, r


def test_single_fixed_point_is_dominant_eigenpair(single_case):
    ops, g, E, r = single_case
    L = r["L"]
    h = r["sol"][1]
    h = h / (ops.u @ h)
    lam = ops.u @ (L @ h)
    assert np.linalg.norm(L @ h - lam * h) <= 1e-8 * np.linalg.norm(L @ h)
    K = ops.C + ops.Araw * (E / g.vth) * (op.QE / op.ME)
    ev, V = scipy.linalg.eig(K, ops.M)
    ev = ev[np.isfinite(ev)]
    k = np.argmin(np.abs(ev - lam))
    assert abs(ev[k] - lam) <= 1e-8 * np.abs(ev).max()
    # the steady EDF is the mode that dominates in time: largest real part
    assert ev[k].real >= ev.real.max() - 1e-8 * np.abs(ev).max()


def test_batched_fixed_point_relation_both_jacobians():
    ops, g, procs = oracle_batched(32, 2)
    prm = instance_params(3, seed=5)
    uh = ops.u / np.linalg.norm(ops.u)
    x, q, info = s0.steady_state_solve_each(ops, prm, ops.initial(3), 0.0, 1e-13, 40, jacobian="exact")
    assert all(i["rel"][0] < 1e-13 for i in info)
    for i in range(3):
        n0, Tg, E = prm["n0"][i], prm["Tg"][i], prm["E"][i]
        ns = prm["ns_by_n0"][:, i]
        K = sum(n0 * ns[s] * ops.FOp[s] for s in range(len(ops.FOp))) + n0 * Tg * ops.FOp_g \
            + E * ops.adv_raw * (1.0 / g.vth) * (op.QE / op.ME)
        xi = x[:, i] / (ops.u @ x[:, i])
        mu = n0 * (ns @ ops.W) @ xi
        r = ops.Minv @ (K @ xi) - mu * xi
        c = uh @ r
        # the dense re-evaluation M^-1 (K x) - mu x cancels to ~1e-12 of |M^-1 K x| (measured
        # 3e-12 at 35 Td), so that is its own floor
        assert np.linalg.norm(r - c * uh) <= 1e-11 * np.linalg.norm(ops.Minv @ (K @ xi))
    # the reference (Picard) Jacobian shares the residual, so its Newton step vanishes
    # at the same fixed point
    res, jac = s0.batched_res_jac(ops, prm, "reference")
    xn = x / (ops.u @ x)
    J = jac(xn)
    r = res(xn)
    for i in range(3):
        step = ops.Q @ np.linalg.solve(J[i], -r[:, i])
        assert np.linalg.norm(step) <= 1e-10 * np.linalg.norm(xn[:, i])


def test_refined_fixed_point():
    """refine_fixed_point: (a) its long-double residual K x - mu M x - gamma M u is at the
    rounding of x (no FP64 cancellation floor); (b) the bulk QoIs agree with the FP64 Newton
    to its own floor, so the polish converges to the same fixed point (a dropped or
    mis-signed term would move them by orders of magnitude more); (c) polishing again
    changes nothing.  Instance 2 is at 1.1 Td, where the ionisation rate is ~6e-12 of the
    elastic one and the FP64 Newton's floor shows."""
    ops, g, procs = oracle_batched(32, 2)
    prm = instance_params(3, seed=5)
    prm = {k: v.copy() for k, v in prm.items()}
    prm["E"][2] = 1.1 * prm["n0"][2] * 1e-21
    x, q, info = s0.steady_state_solve_each(ops, prm, ops.initial(3), 0.0, 1e-15, 40, jacobian="exact")
    h, qr = s0.refine_fixed_point(ops, prm, x)
    LD = np.longdouble
    sc = (1.0 / g.vth) * (op.QE / op.ME)
    for i in range(3):
        n0, Tg, E = prm["n0"][i], prm["Tg"][i], prm["E"][i]
        ns = n0 * prm["ns_by_n0"][:, i]
        KL = sum(LD(ns[s]) * ops.FOp[s].astype(LD) for s in range(len(ops.FOp))) \
            + LD(n0 * Tg) * ops.FOp_g.astype(LD) + LD(E * sc) * ops.adv_raw.astype(LD)
        xi = h[:, i].astype(LD) / (ops.u.astype(LD) @ h[:, i].astype(LD))
        wl = (ns[:, None] * ops.W).sum(axis=0).astype(LD)
        F = KL @ xi - (wl @ xi) * (ops.M.astype(LD) @ xi)
        Mu = ops.M.astype(LD) @ ops.u.astype(LD)
        F = F - (Mu @ F) / (Mu @ Mu) * Mu            # gamma M u is free
        scale = np.abs(KL).max() * np.abs(xi).max()
        assert float(np.abs(F).max() / scale) < 1e-17
        for k in ("energy", "mobility", "diffusion"):
            assert abs(qr[k][i] / q[k][i] - 1) < 1e-12, k
        assert abs(qr["rates"][0, i] / q["rates"][0, i] - 1) < 1e-12
    _, q2 = s0.refine_fixed_point(ops, prm, h, iters=2)
    np.testing.assert_allclose(q2["rates"], qr["rates"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("Tg", [1000.0, 3000.0])
def test_elastic_relaxes_to_gas_maxwellian_and_einstein_relation(Tg):
    el = [p for p in oracle_procs() if p.type == "ELASTIC"]
    kT_ev = Tg / op.TEMP_K_1EV
    g = op.grid_single(96, 1, kT_ev, 30 * kT_ev, 5)      # resolves the 0.25 eV cross-section dip
    ops = s0.SingleOps(g, el, 3.22e22, Tg)
    E = 1e-4 * 1e-21 * 3.22e22          # 1e-4 Td: field heating ~1e-5 of k Tg
    # the l = 1 part is ~1e-5 of h here: iterate well below the default atol = 1e-8
    r = s0.single_steady_state(ops, E, atol=1e-16, rtol=1e-16, max_iter=12)
    q = s0.single_qois(ops, r["sol"][1], E)
    assert q["energy"][0] == pytest.approx(1.5 * kT_ev, rel=1e-4)
    assert q["mobility"][0] > 0
    assert q["mobility"][0] / q["diffusion"][0] == pytest.approx(1.0 / kT_ev, rel=2e-4)


def test_batched_qois_match_single_definitions():
    """The batched and single QoI formulas (bte_0d3v_batched.py:1163-1202 vs
    bte_0d3v_solver.py:1234-1276) agree on the same normalised distribution."""
    ops, g, procs = oracle_batched(24, 1)
    sops = s0.SingleOps(g, procs, 3.22e22, 300.0)
    prm = instance_params(2, seed=9)
    x, qb, _ = s0.steady_state_solve_each(ops, prm, ops.initial(2), 0.0, 1e-12, 40)
    for i in range(2):
        qs = s0.single_qois(sops, x[:, i], prm["E"][i])
        assert qs["energy"][0] == pytest.approx(qb["energy"][i], rel=1e-12)
        assert qs["mobility"][0] == pytest.approx(qb["mobility"][i], rel=1e-12)
        assert qs["diffusion"][0] == pytest.approx(qb["diffusion"][i], rel=1e-12)
        for k in range(len(procs)):
            assert qs["rates"][k][0] == pytest.approx(qb["rates"][k, i], rel=1e-12, abs=1e-30)
