"""GPU parity of the e-e collision path (bz_ee.hip, boltzmann_amd/bte_ee.py) against
oracle/ee.py and the oracle's batched Newton with e-e (bte_0d3v_batched.py:444-476,
652-661, 702-743, 882-979 restated).

* assembly: the potential operators (l blocks of hl_op, gl_op) and the banded tensors
  cc_a, cc_b against the oracle's dense ones (and the dense tensors vanish outside the band);
* steady state with e-e (ionisation degrees 1e-2 .. 1e-4), warm-started from the solution
  without e-e as the reference driver does: QoIs within 1e-8 of the oracle's converged
  fixed point.
"""
import numpy as np
import pytest
import torch

from helpers import instance_params, oracle_batched, product_grid, product_xsecs, qoi_errors, rel_err
from oracle import ee as eeo
from oracle import solver0d as s0

pytestmark = pytest.mark.gpu


def _d(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.fixture(scope="module")
def setup():
    from boltzmann_amd.bte_ee import EEOperator
    from boltzmann_amd.solver0d import Batched0D
    print("case: e-e assembly Nr=24 lmax=1", flush=True)
    ops, g, procs = oracle_batched(24, 1)
    vg = product_grid(g)
    s = Batched0D(vg, product_xsecs(), ee_pattern=True)
    eop = EEOperator(vg, s)
    torch.cuda.synchronize()
    hl, gl = eeo.rosenbluth_ops(g, ops.Minv)
    A, Bt = eeo.cc_tensors(g)
    return ops, g, s, eop, hl, gl, A, Bt


def test_potential_operators(setup):
    ops, g, s, eop, hl, gl, A, Bt = setup
    ns = g.num_sh
    for l in range(ns):
        for dev_op, ref in ((eop.H, hl), (eop.G, gl)):
            r = ref[l::ns, l::ns]
            got = dev_op[l].cpu().numpy()
            assert np.abs(got - r).max() <= 1e-11 * np.abs(r).max()
    # block diagonal in l
    for l in range(ns):
        for m in range(ns):
            if l != m:
                assert np.abs(hl[l::ns, m::ns]).max() == 0.0


def test_banded_tensors(setup):
    ops, g, s, eop, hl, gl, A, Bt = setup
    ns, npr = g.num_sh, g.num_p
    ccA = eop.ccA.cpu().numpy().reshape(npr, 7, 7, ns, ns, ns)
    ccB = eop.ccB.cpu().numpy().reshape(npr, 7, 7, ns, ns, ns)
    for dense, band in ((A, ccA), (Bt, ccB)):
        D = dense.reshape(npr, ns, npr, ns, npr, ns)
        scale = np.abs(D).max()
        inside = np.zeros_like(D, dtype=bool)
        for p in range(npr):
            for dk in range(-3, 4):
                for dr in range(-3, 4):
                    k, r = p + dk, p + dr
                    if 0 <= k < npr and 0 <= r < npr:
                        inside[p, :, k, :, r, :] = True
                        ref = D[p, :, k, :, r, :]                       # (q, l, s)
                        got = band[p, dk + 3, dr + 3]
                        assert np.abs(got - ref).max() <= 1e-12 * scale, (p, dk, dr)
        assert np.abs(D[~inside]).max() == 0.0


def test_steady_state_with_ee(setup):
    ops, g, s, eop, hl, gl, A, Bt = setup
    print("case: e-e steady state Nr=24 lmax=1 B=3", flush=True)
    B = 3
    prm = instance_params(B, seed=3)
    prm["ne"] = prm["n0"] * np.array([1e-2, 1e-3, 1e-4])
    # warm start without e-e (bte_0d3v_batched_driver.py:192-204)
    x0, q0, _ = s0.steady_state_solve_each(ops, prm, ops.initial(B), 0.0, 1e-13, 40, jacobian="exact")
    ops.ee = {"cc": eeo.ee_operator(ops.Minv, A, Bt, hl, gl)}
    try:
        # the oracle's Newton with the reference's e-e Jacobian (cc x + cc_12 x) plus the
        # growth derivative, to the FP64 floor: the fixed point both solvers approach
        x_ref, q_ref, info = s0.steady_state_solve_each(ops, prm, x0, 0.0, 1e-14, 60, jacobian="exact")
    finally:
        ops.ee = None
    x = _d(x0)
    n0, Tg, E, ns, ne = (_d(prm[k]) for k in ("n0", "Tg", "E", "ns_by_n0", "ne"))
    # solved to the FP64 floor: the tail rates need |F| ~ 1e-15 |F0| (a residual of 4e-14 |F0|
    # left 1e-8 on a rate at 2e-6 of the largest)
    out = eop.steady_state(x, n0, Tg, E, ns, ne, 0.0, 1e-15, 60)
    torch.cuda.synchronize()
    st, ra, r0 = (out[k].cpu().numpy() for k in ("status", "res_abs", "res0"))
    print("newton sweeps", out["outer"], "steps", out["iters"].cpu().numpy(), "krylov",
          out["krylov"].cpu().numpy(), "status", st, "|F|/|F0|", ra / r0, flush=True)
    # converged, or stopped at the FP64 evaluation floor of |F| (measured 1e-14 |F0|)
    assert np.all((st == 0) | ((st == 3) & (ra <= 1e-12 * r0)))
    h, q = s.qois(x, n0, E)
    qg = {k: v.cpu().numpy() for k, v in q.items()}
    errs = qoi_errors(qg, q_ref)
    print("qoi errors (rate floor 1e-9)", errs, flush=True)
    # against the oracle's converged fixed point (exact Newton to 1e-14) every QoI meets 1e-8
    # (measured: energy 6.4e-11, rates 9.5e-10, the spread of the operator assembly: the
    # oracle's fixed point with this library's H, G, cc_a, cc_b evaluated in long double is
    # 9.7e-10 away on the same rate)
    assert max(errs[k] for k in ("energy", "mobility", "diffusion")) < 1e-9, errs
    assert errs["rates"] < 2e-8, errs
    hg = h.cpu().numpy()
    for i in range(B):
        assert rel_err(hg[:, i], x_ref[:, i]) < 1e-8


def test_newton_krylov_pieces(setup):
    """Residual, Jacobian-vector product and preconditioner of the e-e Newton-Krylov solve
    against the oracle's dense residual and Jacobian (cc x + cc_12 x, growth derivative)."""
    from boltzmann_amd import _lib
    from boltzmann_amd.bte_ee import _NKWork
    ops, g, s, eop, hl, gl, A, Bt = setup
    print("case: e-e Newton-Krylov pieces Nr=24 lmax=1 B=3", flush=True)
    B = 3
    n = g.num_p * g.num_sh
    prm = instance_params(B, seed=5)
    prm["ne"] = prm["n0"] * np.array([1e-2, 1e-3, 1e-4])
    x0, _, _ = s0.steady_state_solve_each(ops, prm, ops.initial(B), 0.0, 1e-13, 40, jacobian="exact")
    x0 = x0 / (ops.u @ x0)[None, :]
    cc = eeo.ee_operator(ops.Minv, A, Bt, hl, gl)
    ops.ee = {"cc": cc}
    try:
        res, jac = s0.batched_res_jac(ops, prm, jacobian="exact")
        r_ref = res(x0)
        J_ref = jac(x0)
    finally:
        ops.ee = None
    Q = ops.Q
    x = _d(x0)
    n0, Tg, E, ns, ne = (_d(prm[k]) for k in ("n0", "Tg", "E", "ns_by_n0", "ne"))
    w = _NKWork(n, B, 4, x.device, s.tables.P)
    h, gg, negam = eop.potentials(x, ne)
    ee = _lib.EE0D(ns=eop.ns, npr=eop.npr, ccA=eop.ccA.data_ptr(), ccB=eop.ccB.data_ptr(), h=h.data_ptr(),
                   g=gg.data_ptr(), negam=negam.data_ptr())
    bt, _ = s.batch_struct(x, n0, Tg, E, ns)
    grid = s.tables.struct
    _lib.call("bz_ee_newton_residual", grid, bt, ee, w.mu, w.F, w.fn)
    F = w.F.cpu().numpy().T
    # the e-e term is evaluated in two association orders (dense M^-1 cc on the oracle side,
    # banded M-space sums then M^-1 here): agreement to ~1e-8 of |F| at the warm start
    scale = np.abs(r_ref).max(axis=0)
    print("residual err", np.abs(Q.T @ F - r_ref).max(axis=0) / scale, flush=True)
    assert np.all(np.abs(Q.T @ F - r_ref).max(axis=0) <= 1e-7 * scale)
    np.testing.assert_allclose(w.fn.cpu().numpy(), np.linalg.norm(r_ref, axis=0), rtol=1e-7)
    # J v for v in u-perp
    rng = np.random.default_rng(7)
    y = rng.standard_normal((n - 1, B))
    v = Q @ y
    w.z.copy_(_d(v.T))
    w.zT.copy_(_d(v))
    _lib.call("bz_ee_potentials_cols", eop.H, eop.G, eop.ns, eop.npr, w.zT, B, w.hz, w.gz)
    _lib.call("bz_ee_jacobian_t2", eop.ns, eop.npr, B, eop.ccA, eop.ccB, x, w.hz, w.gz, negam, None, w.t2)
    _lib.call("bz_ee_newton_jv", grid, bt, None, w.z, w.t2, w.Jz)
    jv = Q.T @ w.Jz.cpu().numpy().T
    jv_ref = np.einsum("xab,bx->ax", J_ref, y)
    err = np.abs(jv - jv_ref).max(axis=0) / np.abs(jv_ref).max(axis=0)
    print("J v err", err, flush=True)
    assert np.all(err <= 1e-7)
    # preconditioner: z in u-perp with Q^T (L_x - mu) z = Q^T v
    _lib.call("bz_ee_newton_factor", grid, bt, None, w.lprof, w.z1, w.s1)
    _lib.call("bz_ee_newton_precond_apply", grid, bt, None, w.lprof, w.z1, w.s1, w.z, w.Jz, w.zT)
    z = w.Jz.cpu().numpy().T
    nsb = prm["n0"] * prm["ns_by_n0"]
    gam = eeo.gamma_a(ops, x0, prm["ne"])
    for b in range(B):
        L = (np.einsum("s,suv->uv", nsb[:, b], ops.c_en) + prm["n0"][b] * prm["Tg"][b] * ops.c_gT
             + prm["E"][b] * ops.adv + prm["ne"][b] * gam[b] * np.tensordot(cc, x0[:, b], axes=([2], [0])))
        mu = nsb[:, b] @ (ops.W @ x0[:, b])
        lhs = Q.T @ (L @ z[:, b] - mu * z[:, b])
        assert abs(ops.u @ z[:, b]) <= 1e-12 * np.abs(ops.u).max() * np.abs(z[:, b]).max()
        e = np.abs(lhs - y[:, b]).max() / np.abs(y[:, b]).max()
        print("precond err", b, e, flush=True)
        # a solve with K_x - mu M (condition ~1e8 here); only the Krylov convergence rate
        # depends on it
        assert e <= 1e-7
    np.testing.assert_allclose(w.zT.cpu().numpy(), z, rtol=0, atol=1e-14 * np.abs(z).max())


def test_steady_state_with_ee_lmax4():
    """e-e at lmax 4 (window 20 of the solver, 125 harmonic triples) against the oracle."""
    from boltzmann_amd.bte_ee import EEOperator
    from boltzmann_amd.solver0d import Batched0D
    print("case: e-e steady state Nr=12 lmax=4 B=2", flush=True)
    ops, g, procs = oracle_batched(12, 4)
    vg = product_grid(g)
    s = Batched0D(vg, product_xsecs(), ee_pattern=True)
    eop = EEOperator(vg, s)
    B = 2
    prm = instance_params(B, seed=21)
    prm["ne"] = prm["n0"] * np.array([1e-2, 1e-4])
    x0, _, _ = s0.steady_state_solve_each(ops, prm, ops.initial(B), 0.0, 1e-13, 40, jacobian="exact")
    hl, gl = eeo.rosenbluth_ops(g, ops.Minv)
    A, Bt = eeo.cc_tensors(g)
    ops.ee = {"cc": eeo.ee_operator(ops.Minv, A, Bt, hl, gl)}
    try:
        _, q_ref, _ = s0.steady_state_solve_each(ops, prm, x0, 0.0, 1e-14, 60, jacobian="exact")
    finally:
        ops.ee = None
    x = _d(x0)
    n0, Tg, E, ns, ne = (_d(prm[k]) for k in ("n0", "Tg", "E", "ns_by_n0", "ne"))
    out = eop.steady_state(x, n0, Tg, E, ns, ne, 0.0, 1e-15, 60)
    st, ra, r0 = (out[k].cpu().numpy() for k in ("status", "res_abs", "res0"))
    print("status", st, "|F|/|F0|", ra / r0, "steps", out["iters"].cpu().numpy(), flush=True)
    assert np.all((st == 0) | ((st == 3) & (ra <= 1e-12 * r0)))
    h, q = s.qois(x, n0, E)
    errs = qoi_errors({k: v.cpu().numpy() for k, v in q.items()}, q_ref)
    print("qoi errors", errs, flush=True)
    assert max(errs[k] for k in ("energy", "mobility", "diffusion")) < 1e-9, errs
    assert errs["rates"] < 2e-8, errs


def test_steady_state_with_ee_n387():
    """e-e at the size of the batched e-e workload (Nr = 128, lmax = 2, n = 387) against the
    oracle's dense exact Newton, stored in tests/golden/ee_n387.npz (make_ee387.py).

    The assembly and the solve are checked separately.
    * Assembly: H, G and the banded tensors against the oracle's (entrywise the tensors agree
      to 2.7e-16 of their largest entry, H and G to 1e-14: H = M_r^-1 raw with cond(M_r) =
      5.8e6 carries the rounding of raw, which agrees to 1e-15).
    * Solve: with the oracle's H, G and tensors uploaded (K and M stay this library's),
      energy, mobility and diffusion match to 1e-9 and the rates to 2e-8 (measured 1.7e-10
      and 1.6e-8).
    * With this library's operators the most e-e dominated instance (ionisation degree 1e-2)
      differs by 2.0e-7 on its ionisation rate (energy 2.0e-9); degrees 1e-3, 1e-4, 1e-6 by
      6.0e-8, 1.3e-8 and 1e-11.  That is this fixed point's FP64 conditioning, not an error
      of either side: the oracle alone, with its tensors multiplied entrywise by
      1 + 1e-16 N(0, 1), moves by 2.4e-7 on that rate (energy 2.8e-9).  The oracle's H, G
      with this library's tensors change the result by less than 1e-9, so the tensors' last
      bits, which no two FP64 assemblies share, set it.
    """
    import os
    from boltzmann_amd.bte_ee import EEOperator
    from boltzmann_amd.solver0d import Batched0D
    print("case: e-e steady state Nr=128 lmax=2 B=4 vs stored oracle", flush=True)
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "ee_n387.npz"))
    ops, g, procs = oracle_batched(128, 2)
    vg = product_grid(g)
    s = Batched0D(vg, product_xsecs(), ee_pattern=True)
    eop = EEOperator(vg, s)
    ns_, npr = g.num_sh, g.num_p
    hl, gl = eeo.rosenbluth_ops(g, ops.Minv)
    A, Bt = eeo.cc_tensors(g)
    # assembly parity
    for l in range(ns_):
        for dev_op, r_ in ((eop.H, hl), (eop.G, gl)):
            r = r_[l::ns_, l::ns_]
            assert np.abs(dev_op[l].cpu().numpy() - r).max() <= 1e-12 * np.abs(r).max()
    bands = []
    for dense, band in ((A, eop.ccA), (Bt, eop.ccB)):
        D = dense.reshape(npr, ns_, npr, ns_, npr, ns_)
        out = np.zeros((npr, 7, 7, ns_, ns_, ns_))
        for p in range(npr):
            for dk in range(-3, 4):
                for dr in range(-3, 4):
                    k, r = p + dk, p + dr
                    if 0 <= k < npr and 0 <= r < npr:
                        out[p, dk + 3, dr + 3] = D[p, :, k, :, r, :]
        got = band.cpu().numpy().reshape(out.shape)
        assert np.abs(got - out).max() <= 1e-12 * np.abs(out).max()
        bands.append(out.reshape(band.shape))
    del A, Bt
    n0, Tg, E, ns, ne = (_d(ref[k]) for k in ("n0", "Tg", "E", "ns_by_n0", "ne"))
    q_ref = {k: ref[k] for k in ("energy", "mobility", "diffusion", "rates")}

    def solve():
        x = s.initial(4)
        s.solve_raw(x, n0, Tg, E, ns, atol=0.0, rtol=1e-14, max_iter=40, mode=1)
        out = eop.steady_state(x, n0, Tg, E, ns, ne, 0.0, 1e-15, 60)
        st, ra, r0 = (out[k].cpu().numpy() for k in ("status", "res_abs", "res0"))
        print("status", st, "|F|/|F0|", ra / r0, "steps", out["iters"].cpu().numpy(), flush=True)
        assert np.all((st == 0) | ((st == 3) & (ra <= 1e-12 * r0)))
        h, q = s.qois(x, n0, E)
        qn = {k: v.cpu().numpy() for k, v in q.items()}
        return qoi_errors(qn, q_ref), qn

    def per_instance(q):
        return [qoi_errors({k: (v[..., i:i + 1]) for k, v in q.items()},
                           {k: (v[..., i:i + 1]) for k, v in q_ref.items()}) for i in range(4)]

    errs_own, q_own = solve()
    print("qoi errors, this library's operators", errs_own, flush=True)
    print("  per instance (ion deg 1e-2, 1e-3, 1e-4, 1e-6)", per_instance(q_own), flush=True)
    # the oracle's H, G with this library's tensors, then all of the oracle's operators
    for l in range(ns_):
        eop.H[l].copy_(_d(hl[l::ns_, l::ns_]))
        eop.G[l].copy_(_d(gl[l::ns_, l::ns_]))
    errs_hg, q_hg = solve()
    print("qoi errors, the oracle's H, G", errs_hg, per_instance(q_hg), flush=True)
    eop.ccA.copy_(_d(bands[0]))
    eop.ccB.copy_(_d(bands[1]))
    errs, q_all = solve()
    print("qoi errors, the oracle's operators", errs, per_instance(q_all), flush=True)
    assert max(errs_own[k] for k in ("energy", "mobility", "diffusion")) < 1e-8, errs_own
    assert errs_own["rates"] < 5e-7, errs_own
    own = per_instance(q_own)
    assert own[2]["rates"] < 5e-8 and max(own[3].values()) < 1e-8, own
    assert max(errs[k] for k in ("energy", "mobility", "diffusion")) < 1e-9, errs
    assert errs["rates"] < 2e-8, errs
