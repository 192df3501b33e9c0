"""Oracle: 0D steady-state solvers (single and batched), QoIs (FP64, numpy).

Restates (BESolver/python/...):
  * bte_0d3v_batched.assemble_operators       bte_0d3v_batched.py:310-479 (no e-e here)
  * get_rhs_and_jacobian (no e-e branch)      bte_0d3v_batched.py:602-646, 763-821
  * steady_state_solve (Armijo Newton)        bte_0d3v_batched.py:882-979
  * normalized_distribution / compute_QoIs    bte_0d3v_batched.py:1163-1217
  * refine_fixed_point: the same fixed point polished with a long-double
    residual (a parity target, not a reference algorithm)
  * single path _QRQt, newton_solve,
    steady_state_solver, compute_QoIs        bte_0d3v_solver.py:404-460, 24-73, 741-859, 1234-1276

  * newton_solver_batched, step, solve("transient")
                                             bte_0d3v_batched.py:59-126, 981-1161

Pinned by tests/test_oracle_solver.py: the single-path fixed point equals the
generalized eigenvector of (C + E A, M) with u^T h = 1 (scipy.linalg.eig);
the batched fixed point satisfies the same eigen relation up to the
reference's boundary-flux term (W excludes advection); E = 0 elastic-only
gives the Maxwellian at Tg; Maxwellian closed forms pin the QoI vectors;
refine_fixed_point's result has a long-double residual at the rounding of x and
matches the FP64 Newton where that is well conditioned.
"""
import math
import numpy as np

from . import operators as op


# process types that change the number of electrons (they alone carry a growth term)
NONCONSERVATIVE = ("IONIZATION", "ATTACHMENT")


class BatchedOps:
    """Per-grid operators of the batched path (bte_0d3v_batched.py:360-442)."""

    def __init__(self, g: op.Grid, procs, ee=None):
        self.g = g
        self.procs = procs
        species = []
        for p in procs:
            if p.species not in species:
                species.append(p.species)
        self.species = species
        n = g.n
        M = op.mass_matrix(g)
        Minv = op.inverse_mass_matrix(M)
        FOp = [np.zeros((n, n)) for _ in species]
        FOp_grow = [np.zeros((n, n)) for _ in species]
        FOp_g = np.zeros((n, n))
        for p in procs:
            Cp = op.collision_matrix(g, p, tgK=0.0)
            FOp[species.index(p.species)] += Cp
            if p.type in NONCONSERVATIVE:
                FOp_grow[species.index(p.species)] += Cp
            if p.type == "ELASTIC":
                FOp_g += op.gas_temperature_matrix(g, p)
        self.M = M
        self.Minv = Minv
        self.FOp = FOp                 # un-premultiplied collision operators, per species
        self.FOp_g = FOp_g
        self.adv_raw = op.advection_matrix(g)
        self.c_en = np.array([Minv @ F for F in FOp])
        self.c_gT = Minv @ FOp_g
        self.adv = (Minv @ self.adv_raw) * (1.0 / g.vth) * (op.QE / op.ME)
        self.mass = op.mass_op(g)
        self.temp = op.temp_op(g) * op.TEMP_EV_FAC
        self.mobility = op.mobility_op(g)
        self.diffusion = op.diffusion_op(g, procs)
        self.rates = np.array([op.rate_op(g, p) for p in procs])
        mw0 = op.maxwellian0(g.vth)
        self.u = self.mass * mw0 * g.vth ** 3
        self.Q, self.R = qr_basis(self.u, g.num_p, g.num_sh)
        # growth vectors W_s = u^T M^-1 C_s (bte_0d3v_batched.py:636) from the processes that
        # change the electron number; for elastic and excitation processes u^T M^-1 C is zero
        # (particle conservation, checked by the oracle's tests) and the reference's product
        # leaves only roundoff there, ~1e-17 of the ionisation part, which moves rates at
        # 1e-11 of the largest by ~1e-6 (see test_oracle_solver.py)
        self.W = np.einsum("v,svu->su", self.u, np.array([Minv @ F for F in FOp_grow]))
        self.ee = ee

    def initial(self, B):
        h = op.maxwellian_init(self.g, self.Minv)
        return np.repeat(h.reshape(-1, 1), B, axis=1)


def qr_basis(u, num_p, num_sh):
    """bte_0d3v_batched.py:425-442: QR of I - p p^T, drop column/row (num_p-1)*num_sh."""
    n = u.shape[0]
    pv = u.reshape((n, 1)) / np.sqrt(np.dot(u, u))
    Impp = np.eye(n) - pv @ pv.T
    Qm, Rm = np.linalg.qr(Impp)
    k = (num_p - 1) * num_sh + 0
    Q = np.delete(Qm, k, axis=1)
    R = np.delete(Rm, k, axis=0)
    assert np.linalg.norm(Impp - Q @ R) / np.linalg.norm(Impp) < 1e-10
    assert np.linalg.norm(Q.T @ Q - np.eye(n - 1)) < 1e-10
    return Q, R


def batched_res_jac(ops: BatchedOps, params, jacobian="reference"):
    """get_rhs_and_jacobian, no e-e branch (bte_0d3v_batched.py:763-821).

    params: dict with n0 (B,), Tg (B,), E (B,), ns_by_n0 (S,B) [ne (B,) for e-e].
    jacobian="exact" adds the derivative of the growth term, -Q^T x (w Q), which the
    reference Jacobian omits; both share the residual and hence the fixed point."""
    Q, QT = ops.Q, ops.Q.T
    QT_Cen = np.einsum("mu,suv->smv", QT, ops.c_en)
    QT_Cgt = QT @ ops.c_gT
    QT_A = QT @ ops.adv
    QT_Cen_Q = np.einsum("av,svu->sau", QT, ops.c_en @ Q)
    QT_Cgt_Q = QT @ ops.c_gT @ Q
    QT_A_Q = QT @ ops.adv @ Q
    W = ops.W
    n0, Tg, E, ns_by_n0 = params["n0"], params["Tg"], params["E"], params["ns_by_n0"]
    I = np.eye(QT.shape[0])

    ee = ops.ee
    if ee is not None:
        from . import ee as eeo
        cc = ee["cc"]
        cc12 = np.swapaxes(cc, 1, 2)
        ne = params["ne"]

    def res(x):
        y = (n0 * (np.einsum("sx,svx->vx", ns_by_n0, np.einsum("svu,ux->svx", QT_Cen, x))
                   + Tg * (QT_Cgt @ x))
             + E * (QT_A @ x)
             - n0 * np.einsum("sx,sx->x", ns_by_n0, W @ x) * (QT @ x))
        if ee is not None:
            # + ne gamma_a(x) Q^T (cc x) x   (bte_0d3v_batched.py:702-709)
            y = y + ne * eeo.gamma_a(ops, x, ne) * (QT @ eeo.ee_term(cc, x))
        return y

    def jac(x):
        ns = n0 * ns_by_n0
        L = np.einsum("sx,svu->xvu", ns, QT_Cen_Q) + np.einsum("a,bc->abc", n0 * Tg, QT_Cgt_Q)
        mu = np.einsum("sx,sx->x", ns, W @ x)
        J = L + np.einsum("a,bc->abc", E, QT_A_Q) - np.einsum("a,bc->abc", mu, I)
        if ee is not None:
            # + ne gamma_a Q^T (cc x + cc_12 x) Q   (bte_0d3v_batched.py:734-743)
            c1 = np.tensordot(cc, x, axes=([2], [0])) + np.tensordot(cc12, x, axes=([2], [0]))   # (n, n, B)
            QTcQ = np.einsum("bi,ijx,jc->xbc", QT, c1, Q)
            J = J + np.einsum("x,xbc->xbc", ne * eeo.gamma_a(ops, x, ne), QTcQ)
        if jacobian == "exact":
            wq = np.einsum("sx,su,uk->xk", ns, W, Q)
            J = J - np.einsum("vx,xk->xvk", QT @ x, wq)
        return J

    return res, jac


def steady_state_solve(ops: BatchedOps, params, f0, atol, rtol, max_iter,
                       alpha_min=1e-16, alpha_rho=0.5, line_search_c=1e-4,
                       res_jac=None, return_info=False, jacobian="reference"):
    """bte_0d3v_batched.py:882-979 (batch-coupled loop exactly as written)."""
    u = ops.u
    Q, R = ops.Q, ops.R
    res_func, jac_func = res_jac if res_jac is not None else batched_res_jac(ops, params, jacobian)
    n_pts = f0.shape[1]
    f0 = f0 / (u @ f0)
    fb_prev = R @ f0
    f1 = u / np.dot(u, u)
    h_prev = f1[:, None] + Q @ fb_prev
    r0 = np.linalg.norm(res_func(h_prev), axis=0)
    abs_error = r0.copy()
    it = 0
    h_curr = h_prev
    diverged = False
    while (abs_error > atol + rtol * r0).any() and it < max_iter:
        L = jac_func(h_prev)
        rhs = res_func(h_prev)
        abs_error = np.linalg.norm(rhs, axis=0)
        Linv = np.linalg.inv(L)
        pp_mat = np.einsum("abc,ca->ba", Linv, -rhs)
        p = Q @ pp_mat
        alpha = np.ones(n_pts)
        rf_new = 0.5 * np.linalg.norm(res_func(h_prev + alpha * p), axis=0) ** 2
        while (rf_new > 0.5 * abs_error ** 2 - line_search_c * alpha * abs_error ** 2).any():
            rc = rf_new > 0.5 * abs_error ** 2 - line_search_c * alpha * abs_error ** 2
            alpha[rc] *= alpha_rho
            if (alpha < alpha_min).all():
                diverged = True
                break
            rf_new = 0.5 * np.linalg.norm(res_func(h_prev + alpha * p), axis=0) ** 2
        if diverged:
            break
        h_curr = h_prev + alpha * p
        if it % 10 == 0:
            abs_error = np.linalg.norm(res_func(h_curr), axis=0)
        h_prev = h_curr
        it += 1
    final_abs = np.linalg.norm(res_func(h_curr), axis=0)
    h = normalized_distribution(ops, h_curr)
    qoi = compute_qois(ops, params, h)
    if return_info:
        return h, qoi, {"iters": it, "abs": final_abs, "rel": final_abs / r0, "diverged": diverged}
    return h, qoi


def refine_fixed_point(ops: BatchedOps, params, h0, iters=4):
    """The fixed point of the batched residual (get_rhs_and_jacobian, bte_0d3v_batched.py:
    763-821) polished to long-double accuracy, for use as a parity target.

    Q^T M^-1 (K x - mu(x) M x) = 0 with x = f1 + Q y holds iff K x - mu M x = gamma M u for
    some gamma, with u.x = 1 (the columns of Q span u-perp and M^-1 M u = u).  So this runs
    Newton on the bordered system [K - mu M - M x w^T, -M u; u^T, 0] [dx; dgamma] =
    -[K x - mu M x - gamma M u; u.x - 1], the residual formed in long double from the same
    FP64 operators, the step solved in FP64.  The premultiplied Q-basis residual of the
    reference rounds at ~1e-16 |M^-1| |K| |x|; this one at ~1e-19 |K| |x|, which matters for
    rates 1e-11 below the largest (measured: 3e-8 relative at E/N ~ 1 Td).
    h0: (n, B) unnormalised iterates near the fixed point.  Returns (h, qoi) as
    steady_state_solve does."""
    LD = np.longdouble
    g = ops.g
    n, B = h0.shape
    sc = (1.0 / g.vth) * (op.QE / op.ME)
    u = ops.u
    M = ops.M
    ML, uL = M.astype(LD), u.astype(LD)
    MuL = ML @ uL
    Mu = M @ u
    out = np.empty_like(h0)
    for i in range(B):
        n0, Tg, E = params["n0"][i], params["Tg"][i], params["E"][i]
        ns = n0 * params["ns_by_n0"][:, i]
        KL = sum(LD(ns[s]) * ops.FOp[s].astype(LD) for s in range(len(ops.FOp)))
        KL = KL + LD(n0 * Tg) * ops.FOp_g.astype(LD) + LD(E * sc) * ops.adv_raw.astype(LD)
        K = KL.astype(np.float64)
        wL = (ns[:, None] * ops.W).sum(axis=0).astype(LD)
        w = wL.astype(np.float64)
        x = h0[:, i].astype(LD) / (uL @ h0[:, i].astype(LD))
        gam = LD(0)
        for _ in range(iters):
            xf = x.astype(np.float64)
            mu = wL @ x
            F = KL @ x - mu * (ML @ x) - gam * MuL
            J = np.zeros((n + 1, n + 1))
            J[:n, :n] = K - float(mu) * M - np.outer(M @ xf, w)
            J[:n, n] = -Mu
            J[n, :n] = u
            d = np.linalg.solve(J, -np.concatenate([F.astype(np.float64), [float(uL @ x - 1)]]))
            x = x + d[:n].astype(LD)
            gam = gam + LD(d[n])
        out[:, i] = x.astype(np.float64)
    h = normalized_distribution(ops, out)
    return h, compute_qois(ops, params, h)


def steady_state_solve_each(ops: BatchedOps, params, f0, atol, rtol, max_iter, **kw):
    """The batched loop run on every instance alone (batch of one)."""
    B = f0.shape[1]
    hs, qs, infos = [], [], []
    for i in range(B):
        pi = {k: (v[..., i:i + 1] if isinstance(v, np.ndarray) else v) for k, v in params.items()}
        h, q, info = steady_state_solve(ops, pi, f0[:, i:i + 1], atol, rtol, max_iter,
                                        return_info=True, **kw)
        hs.append(h)
        qs.append(q)
        infos.append(info)
    h = np.concatenate(hs, axis=1)
    qoi = {k: np.concatenate([q[k] for q in qs], axis=-1) for k in qs[0]}
    return h, qoi, infos


def normalized_distribution(ops: BatchedOps, ff):
    """bte_0d3v_batched.py:1204-1217."""
    g = ops.g
    mm_fac = (1.0 / (2.0 * math.sqrt(math.pi))) * 4.0 * math.pi
    scale = (ops.mass / mm_fac) @ ff * (2.0 * (g.vth / op.C_GAMMA) ** 3)
    return ff / scale


def compute_qois(ops: BatchedOps, params, ff, effective_mobility=True):
    """bte_0d3v_batched.py:1163-1202."""
    g = ops.g
    ns = g.num_sh
    mw0 = op.maxwellian0(g.vth)
    vth = g.vth
    c = op.C_GAMMA
    mm = (ops.mass @ ff) * mw0 * vth ** 3
    mu = 1.5 * (ops.temp @ ff) * mw0 * vth ** 5 / mm
    E, n0 = params["E"], params["n0"]
    if effective_mobility:
        M = (ops.mobility @ (math.sqrt(3) * ff[1::ns, :])) * (-(c / (3 * (E / n0))))
    else:
        M = (ops.mobility @ (math.sqrt(3) * ff[1::ns, :])) * (-(c / (3 * (1.0 / n0))))
    D = (ops.diffusion @ ff[0::ns, :]) * (c / 3.0)
    rr = ops.rates @ ff[0::ns, :]
    return {"energy": mu, "mobility": M, "diffusion": D, "rates": rr}


# ----------------------------------------------------------------------------- single path

class SingleOps:
    """bte_0d3v._setup_bsp_sph + _qoi_ops + _QRQt (bte_0d3v_solver.py:161-276, 372-460)."""

    def __init__(self, g: op.Grid, procs, n0, Tg, ns_by_n0=None):
        self.g = g
        self.procs = procs
        species = []
        for p in procs:
            if p.species not in species:
                species.append(p.species)
        if ns_by_n0 is None:
            ns_by_n0 = np.ones(len(species))
        self.n0 = n0
        self.M = op.mass_matrix(g)
        self.Minv = op.inverse_mass_matrix(self.M)
        C = np.zeros((g.n, g.n))
        C_grow = np.zeros((g.n, g.n))
        for p in procs:
            x = ns_by_n0[species.index(p.species)] * n0
            if p.type == "ATTACHMENT":
                Cp = x ** 2 * op.collision_matrix(g, p, tgK=Tg)
            else:
                Cp = x * op.collision_matrix(g, p, tgK=Tg)
            C += Cp
            if p.type in NONCONSERVATIVE:
                C_grow += Cp
        self.C = C
        self.C_grow = C_grow           # the part of C that changes the electron number
        self.Araw = op.advection_matrix(g)
        self.mass = op.mass_op(g)
        self.temp = op.temp_op(g)
        self.mobility = op.mobility_op(g)
        self.diffusion = op.diffusion_op(g, procs)
        self.rates = np.array([op.rate_op(g, p) for p in procs])
        mw0 = op.maxwellian0(g.vth)
        self.u = self.mass * mw0 * g.vth ** 3
        self.Q, self.R = qr_basis(self.u, g.num_p, g.num_sh)


def newton_solve(h0, res_func, jac_func, f1, Q, R, rtol=1e-5, atol=1e-5, max_iter=1000):
    """bte_0d3v_solver.py:24-73."""
    abs_error = 1.0
    rel_error = 1.0
    it = 0
    fb_prev = R @ h0
    h_prev = f1 + Q @ fb_prev
    h_curr = h_prev
    while (rel_error > rtol and abs_error > atol) and it < max_iter:
        L = jac_func(h_prev)
        rhs = -res_func(h_prev)
        abs_error = np.linalg.norm(rhs)
        p = np.linalg.solve(L, rhs)
        p = Q @ p
        alpha = 1.0
        diverged = False
        while np.linalg.norm(res_func(h_prev + alpha * p)) > abs_error:
            alpha *= 0.5
            if alpha < 1e-30:
                diverged = True
                break
        if diverged:
            break
        h_curr = h_prev + alpha * p
        if it % 10 == 0:
            rel_error = np.linalg.norm(h_prev - h_curr) / np.linalg.norm(h_curr)
        h_prev = h_curr
        it += 1
    return h_curr, abs_error, rel_error, it


def single_steady_state(ops: SingleOps, E_field, h_init=None, rtol=1e-13, atol=1e-8, max_iter=300):
    """bte_0d3v_solver.py:741-859 (no e-e)."""
    g = ops.g
    Emat = ops.Araw * (E_field / g.vth) * (op.QE / op.ME)
    CpE = ops.Minv @ (ops.C + Emat)
    u, Q, R = ops.u, ops.Q, ops.R
    QT = Q.T
    # W = u^T M^-1 (C + E A) (bte_0d3v_solver.py:787) without the exact zeros of the
    # particle-conserving processes (see BatchedOps)
    W = u @ (ops.Minv @ (ops.C_grow + Emat))

    def res(x):
        return QT @ (CpE @ x) - (W @ x) * (QT @ x)

    def jac(x):
        return (QT @ CpE - (W @ x) * QT) @ Q

    f1 = u / np.dot(u, u)
    if h_init is None:
        h_init = op.maxwellian_init(g, ops.Minv)
    h_init = h_init / np.dot(u, h_init)
    h, a, r, it = newton_solve(h_init, res, jac, f1, Q, R, rtol=rtol, atol=atol, max_iter=max_iter)
    sol = np.zeros((2, g.n))
    sol[0] = normalized_single(ops, h_init)
    sol[1] = normalized_single(ops, h)
    return {"sol": sol, "atol": a, "rtol": r, "iters": it, "tgrid": None, "L": CpE}


def normalized_single(ops, f):
    """utils.py:980-997."""
    mm_fac = (1.0 / (2.0 * math.sqrt(math.pi))) * 4.0 * math.pi
    scale = np.dot(f, ops.mass / mm_fac) * (2 * (ops.g.vth / op.C_GAMMA) ** 3)
    return f / scale


def single_qois(ops: SingleOps, hh, E_field, effective_mobility=True):
    """bte_0d3v_solver.py:1234-1276."""
    g = ops.g
    ns = g.num_sh
    hh = hh.reshape((-1, g.n))
    mw0 = op.maxwellian0(g.vth)
    vth = g.vth
    c = op.C_GAMMA
    eavg_to_K = 2.0 / (3.0 * op.KB)
    mm = (hh @ ops.mass) * mw0 * vth ** 3
    mu = (hh @ ops.temp) * mw0 * vth ** 5 * (0.5 * op.ME * eavg_to_K / mm) * 1.5 / op.TEMP_K_1EV
    if effective_mobility:
        M = (math.sqrt(3) * hh[:, 1::ns]) @ ops.mobility * (-(c / (3 * (E_field / ops.n0))))
    else:
        M = (math.sqrt(3) * hh[:, 1::ns]) @ ops.mobility * (-(c / (3 * (1.0 / ops.n0))))
    D = (hh[:, 0::ns] @ ops.diffusion) * (c / 3.0)
    rr = [hh[:, 0::ns] @ ops.rates[i] for i in range(len(ops.procs))]
    return {"energy": mu, "mobility": M, "diffusion": D, "rates": rr}


# ----------------------------------------------------------------------------- transient

def newton_solver_batched(x, residual, jac_inv, atol, rtol, iter_max):
    """bte_0d3v_batched.py:59-126: frozen Jacobian inverse, full steps; the convergence
    test of an iterate is made before its update and the update is still applied."""
    r0 = residual(x)
    norm_r0 = np.linalg.norm(r0, axis=0)
    norm_rr = norm_r0
    count = 0
    converged = (norm_rr / norm_r0 < rtol).all() or (norm_rr < atol).all()
    while not converged and count < iter_max:
        rr = residual(x)
        norm_rr = np.linalg.norm(rr, axis=0)
        converged = (norm_rr / norm_r0 < rtol).all() or (norm_rr < atol).all()
        p = np.einsum("ijk,ki->ji", jac_inv, -rr)
        x = x + p
        count += 1
    return {"status": converged, "x": x, "atol": norm_rr, "rtol": norm_rr / norm_r0, "iter": count}


def transient_step(ops: BatchedOps, params, u, dt_tau, atol, rtol, max_iter, du0=None):
    """One backward-Euler step (bte_0d3v_batched.py:981-1021 with dt -> dt * tau as in
    solve("transient"), :1083-1093): u (n, B) normalised, du0 the initial increment in the
    Q basis (zero in step(); the previous step's increment in solve("transient"));
    returns (u_new, info)."""
    res, jac = batched_res_jac(ops, params, "reference")
    Q = ops.Q
    B = u.shape[1]
    I = np.eye(Q.shape[1])
    J = I[None] - dt_tau * jac(u)
    Jinv = np.linalg.inv(J)
    du = np.zeros((Q.shape[1], B)) if du0 is None else du0
    info = newton_solver_batched(du, lambda d: d - dt_tau * res(u + Q @ d), Jinv, atol, rtol, max_iter)
    return u + Q @ info["x"], info


def transient_solve(ops: BatchedOps, params, f0, dt, cycles, Efreq, atol, rtol, max_iter):
    """bte_0d3v_batched.solve(..., "transient") (:1040-1161): E(t) = eRe cos 2 pi t +
    eIm sin 2 pi t (t in RF periods), dt in periods, step dt * tau with tau = 1/Efreq,
    trapezoidal cycle average of the last full cycle; returns (u, qoi of the average)."""
    tau = 1e-7 if Efreq == 0 else 1.0 / Efreq
    io_cycle = 1.0
    io_freq = int(io_cycle / dt)
    steps_total = int(cycles / dt)
    u = f0 / (ops.u @ f0)
    tt = 0.0
    cycle_avg_u = np.zeros_like(u)
    cycle_avg_v = np.zeros_like(u)
    du = np.zeros((ops.Q.shape[1], u.shape[1]))
    for ts_idx in range(steps_total + 1):
        prm = dict(params)
        if Efreq > 0:
            prm["E"] = params["eRe"] * np.cos(2 * np.pi * (tt + dt)) + params["eIm"] * np.sin(2 * np.pi * (tt + dt))
        if ts_idx > 0 and ts_idx % io_freq == 0:
            cycle_avg_u *= 0.5 * dt / io_cycle
            cycle_avg_v[:] = cycle_avg_u
            cycle_avg_u[:] = 0
            if ts_idx == steps_total:
                break
        u_new, info = transient_step(ops, prm, u, dt * tau, atol, rtol, max_iter, du0=du)
        if not info["status"]:
            raise RuntimeError("transient Newton step failed")
        du = info["x"]
        cycle_avg_u += u
        u = u_new
        cycle_avg_u += u
        tt += dt
    h = normalized_distribution(ops, cycle_avg_v)
    q = compute_qois(ops, params, h, effective_mobility=False)
    return u, q
