"""Pins the oracle of the fully implicit 1D3V step on the finite-volume / ordinate grid
(oracle/fvm1d3v.py) on the CPU with what the mathematics and the physics fix: exact
differences of polynomials, the derivative along the field of g(v cos(theta)), the upwind
direction of the advection operator, particle balance of the collision operator against the
rate vectors, the x shift operator on linear functions, each preconditioner sweep as the
exact inverse of its own part of the operator, the GMRES step against the dense solve and
the moments of the Maxwellian initial state against an independent quadrature."""
import math

import numpy as np
import pytest

from helpers import oracle_procs
from oracle import fvm1d3v as fv
from oracle.operators import C_GAMMA


def _solver(Np=10, Nr=15, Nvt=8, q=3, **kw):
    return fv.Fvm1d3v(fv.Params(Np=Np, Nr=Nr, Nvt=Nvt, spline_qpts=q, **kw), oracle_procs())


@pytest.mark.parametrize("n", [2, 3, 4, 5])
def test_fd_weights_exact_on_polynomials(n):
    rng = np.random.default_rng(n)
    x = np.sort(rng.uniform(0.3, 2.0, n))
    for order in (1, 2):
        if order >= n:
            continue
        W = fv.fd_coefficients(x, order)
        for k in range(n):
            f = x ** k
            d = (math.factorial(k) / math.factorial(k - order) * x ** (k - order)) if k >= order else 0 * x
            assert np.abs(W @ f - d).max() <= 1e-9 * max(1.0, np.abs(d).max()), (order, k)


@pytest.mark.parametrize("sw", [2, 3, 4])
def test_upwinded_derivatives(sw):
    """Every row with a stencil differentiates linear functions exactly, rows with a full
    stencil polynomials of degree sw-1; "L" rows use only points at or left of the row, "R"
    rows only points at or right of it (the upwind direction)."""
    x = np.sort(np.random.default_rng(sw).uniform(-1, 1, 17))
    for dir in ("L", "R"):
        D = fv.upwinded_dx(x, 1, sw, dir)
        rows = range(1, 17) if dir == "L" else range(0, 16)
        np.testing.assert_allclose((D @ (3 * x - 1))[list(rows)], 3.0, rtol=1e-11)
        full = range(sw - 1, 17) if dir == "L" else range(0, 17 - sw + 1)
        p = np.polynomial.Polynomial(np.arange(1.0, sw + 1))
        np.testing.assert_allclose((D @ p(x))[list(full)], p.deriv()(x)[list(full)], rtol=1e-9, atol=1e-9)
        nz = np.nonzero(D)
        assert (nz[1] <= nz[0]).all() if dir == "L" else (nz[1] >= nz[0]).all()
    th, _ = fv.vt_grid(16)
    for dir in ("L", "R"):
        D = fv.upwinded_dvt(th, 1, sw, dir)
        rows = slice(1, None) if dir == "L" else slice(0, -1)      # the pole rows have no stencil
        np.testing.assert_allclose((D @ (2 * th + 1))[rows], 2.0, rtol=1e-11)
        nz = np.nonzero(D)
        assert (nz[1] <= nz[0]).all() if dir == "L" else (nz[1] >= nz[0]).all()


def _field_derivative_error(Nvt, Nr=15, q=4):
    """max |A g(v cos) - g'(v cos)| over the rows away from the boundary rows of the
    stencils, for Ep and En, relative to max |g'|."""
    S = _solver(Nr=Nr, Nvt=Nvt, q=q)
    g = lambda z: np.exp(-(z - 0.4) ** 2)
    dg = lambda z: -2 * (z - 0.4) * np.exp(-(z - 0.4) ** 2)
    Ep, En = fv.advection_fvm(S.vr, S.vt, 4, 4, S.domain)
    vz = np.outer(S.vr, np.cos(S.vt))
    f, want = g(vz).reshape(-1), dg(vz)
    Nr, mp = S.Nr, Nvt // 2
    keep = np.zeros((Nr, Nvt), dtype=bool)
    keep[4:Nr - 4] = True
    keep[:, [0, mp - 1, mp, Nvt - 1]] = False
    keep &= S.vr[:, None] < 2.0
    err = 0.0
    for A in (Ep, En):
        err = max(err, np.abs((A @ f).reshape(Nr, Nvt) - want)[keep].max())
    return err / np.abs(want).max()


def test_advection_is_the_derivative_along_the_field():
    """cos(theta) d/dv - (sin(theta)/v) d/dtheta of g(v cos(theta)) is g'(v cos(theta)).  The
    error is led by the low-order angular stencils next to theta = pi/2, where the two
    half-ranges join; it falls about fourfold each time Nvt doubles (a sign or scaling
    mistake in either part leaves an O(1) error that does not fall)."""
    e = [_field_derivative_error(n) for n in (16, 32, 64)]
    assert e[0] / e[1] > 3.0 and e[1] / e[2] > 3.0 and e[2] < 0.015, e


def test_advection_upwind_direction():
    """With E >= 0 the BTE advects f towards -v_z: Ep may couple a point only to points with
    larger v_z (larger v where cos > 0, smaller v where cos < 0, larger ordinate index = smaller
    theta); En the mirror image."""
    S = _solver(Nvt=8)
    Ep_vr, Ep_vt, En_vr, En_vt = fv.advection_fvm_parts(S.vr, S.vt, 4, 4, S.domain)
    Nvt, mp = S.Nvt, S.Nvt // 2
    for A, sgn in ((Ep_vr, 1), (En_vr, -1)):
        i, j = np.nonzero(A)
        ri, ti, rj, tj = i // Nvt, i % Nvt, j // Nvt, j % Nvt
        assert (ti == tj).all()
        up = np.where(ti >= mp, rj >= ri, rj <= ri)
        assert up.all() if sgn > 0 else np.where(ti >= mp, rj <= ri, rj >= ri).all()
    for A, sgn in ((Ep_vt, 1), (En_vt, -1)):
        i, j = np.nonzero(A)
        assert (i // Nvt == j // Nvt).all()
        assert ((j % Nvt) >= (i % Nvt)).all() if sgn > 0 else ((j % Nvt) <= (i % Nvt)).all()


def test_collision_particle_balance():
    """For an isotropic Maxwellian f: elastic and excitation collisions conserve particles
    (mass . C f = 0), ionization creates one electron per event (mass . C f = the ionization
    rate of f), to the accuracy of the linear interpolation at the pre-collision speeds.  The
    rate vectors are the reference's QoI vectors, integrated independently of C."""
    S = _solver(Nr=63, q=5)
    f = S.initial_condition()[:, 3]
    sc = 2 * (S.vth / C_GAMMA) ** 3
    for pr, rate in zip(S.procs, S.op_rate):
        G, dcs, C = fv.collision_fvm(S.vr, S.vr_qw, S.vt_qw, pr, S.vth, S.domain)
        loss = S.op_mass @ (4 * np.pi * np.kron(dcs, np.ones(S.Nvt)) * f)
        assert abs(loss - (rate @ f) / sc) < 1e-10 * loss      # loss integral = the rate
        net = S.op_mass @ (C @ f)
        want = loss if pr.type == "IONIZATION" else 0.0
        assert abs(net - want) < 2e-3 * loss, (pr.type, net, want, loss)
    # a planted error in the ionization Jacobian (2 v_pre / v instead of 4 v_pre / v) must fail
    pr = [p for p in S.procs if p.type == "IONIZATION"][0]
    G, dcs, C = fv.collision_fvm(S.vr, S.vr_qw, S.vt_qw, pr, S.vth, S.domain)
    Cbad = 0.5 * (C + 4 * np.pi * np.kron(np.diag(dcs), np.eye(S.Nvt))) - 4 * np.pi * np.kron(np.diag(dcs), np.eye(S.Nvt))
    loss = S.op_mass @ (4 * np.pi * np.kron(dcs, np.ones(S.Nvt)) * f)
    assert abs(S.op_mass @ (Cbad @ f) - loss) > 0.5 * loss


def test_x_shift_on_linear_functions():
    """The bands are I + dt a_r cos(theta_t) d/dx, upwinded (left neighbours for
    cos >= 0), exact on linear functions except at the inflow row, which is the identity."""
    from oracle.banded import banded_tri_mv
    S = _solver(Np=12, px=2)
    dt = 1e-3
    S.init(dt)
    f = 0.7 * S.xp - 0.2
    for r in (0, S.Nr // 2):
        for t in range(S.Nvt):
            low = S.vt[t] <= 0.5 * np.pi
            d = S.bte_x_shift_diag[r, t][None]
            y = banded_tri_mv(d, f[None], low)[0]
            c = dt * S.op_adv_x_d[r] * np.cos(S.vt[t])
            want = f + c * 0.7
            inflow = 0 if low else S.Nx - 1
            want[inflow] = f[inflow]
            np.testing.assert_allclose(y, want, rtol=1e-12, atol=1e-14)


def _sweep_matrix(fn, n):
    P = np.zeros((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        P[:, j] = fn(e)
        e[j] = 0.0
    return P


def test_preconditioner_sweeps_invert_their_operators():
    """Each XVTVRC stage inverts its own part of the step operator exactly: the x sweep (with
    zeroed inflow) the x shift, the vr and vt sweeps I - dt tau E (q/m)/vth times the radial
    and angular parts of Ep / En, PmatC the collision part."""
    from oracle.banded import adv_vr_BE, adv_vt_BE
    S = _solver(Np=6, Nr=7, Nvt=8, q=2)
    p = S.params
    dt = 2e-3
    S.init(dt)
    Nr, Nvt, Nx = S.Nr, S.Nvt, S.Nx
    # x part alone: no collisions, no field
    S0 = _solver(Np=6, Nr=7, Nvt=8, q=2)
    S0.op_col_en = 0 * S0.op_col_en
    S0.init(dt)
    E0 = np.zeros(Nx)
    v = np.random.default_rng(1).normal(size=(S.dof_v, Nx))
    v = S0.rhs_vector(v).reshape(S.dof_v, Nx)
    y = S0.step_op_split(E0, S0.rhs_mvec(E0, v, dt).reshape(S.dof_v, Nx), dt)
    assert np.abs(y - v).max() < 1e-12 * np.abs(v).max()
    # v parts, one point with E >= 0 and one with E < 0
    Ep_vr, Ep_vt, En_vr, En_vt = fv.advection_fvm_parts(S.vr, S.vt, 4, 4, S.domain)
    for Ex in (3e3, -5e3):
        E = np.array([Ex])
        c = dt * p.tau * Ex * fv.QM / S.vth
        aE = (p.tau * p.qe / p.me / S.vth) * E
        Pvr = _sweep_matrix(lambda b: adv_vr_BE(S.vr_diag_l2r, S.vr_diag_r2l, aE, np.cos(S.vt), dt,
                                                b.reshape(Nr, Nvt).T.reshape(-1), Nr, Nvt, 1).reshape(Nvt, Nr).T.reshape(-1),
                            S.dof_v)
        Pvt = _sweep_matrix(lambda b: adv_vt_BE(S.vt_diag_l2r, S.vt_diag_r2l, aE, S.vr, dt, b, Nr, Nvt, 1).reshape(-1),
                            S.dof_v)
        Avr, Avt = (Ep_vr, Ep_vt) if Ex >= 0 else (En_vr, En_vt)
        I = np.eye(S.dof_v)
        assert np.abs(Pvr @ (I - c * Avr) - I).max() < 1e-11
        assert np.abs(Pvt @ (I + c * Avt) - I).max() < 1e-11
    vmat = p.tau * p.n0 * p.np0 * S.op_col_en
    assert np.abs(S.PmatC @ (np.eye(S.dof_v) - dt * vmat) - np.eye(S.dof_v)).max() < 1e-12


def test_gmres_step_solves_the_dense_system():
    S = _solver(Np=8, Nr=9, Nvt=8, q=3)
    dt = 1e-3
    S.init(dt)
    v = S.initial_condition()
    E = 1e4 * S.xp ** 3
    y = S.step(E, v, dt)
    A = S.rhs_matrix(E, dt)
    yd = np.linalg.solve(A, S.rhs_vector(v)).reshape(S.dof_v, S.Nx)
    assert np.abs(y - yd).max() < 1e-8 * np.abs(yd).max()
    b = np.linalg.norm(S.rhs_vector(v))
    assert S.gmres_info["residual_abs"] < max(S.params.atol, S.params.rtol * b)


def test_maxwellian_state_moments():
    """initial_condition: ne(x) times a Maxwellian at Te: the mass moment returns ne, the
    temperature moment 4 eV up to the truncation of the grid, and the normalised rates are
    <sigma v> of the Maxwellian (scipy quad over energy)."""
    import scipy.constants as sc
    import scipy.integrate as si
    S = _solver(Np=6, Nr=63, q=5)
    v = S.initial_condition()
    q = S.compute_qoi(v)
    xx = S.params.L * (S.xp + 1)
    ne = 1e6 * (1e7 + 1e9 * (1 - 0.5 * xx / S.params.L) ** 2 * (0.5 * xx / S.params.L) ** 2)
    np.testing.assert_allclose(q["ne"], ne, rtol=2e-6)
    assert np.abs(q["Te"] - 4.0).max() < 1e-5
    kT = 4.0
    fe = lambda e: 2.0 * np.sqrt(e / np.pi) * kT ** -1.5 * np.exp(-e / kT)
    vel = lambda e: np.sqrt(2.0 * e * sc.e / sc.m_e)
    for m, pr in enumerate(S.procs):
        sig = lambda e: float(pr.total_cross_section(np.array([e]))[0])
        lo = max(pr.threshold, 1e-8)
        k = si.quad(lambda e: sig(e) * vel(e) * fe(e), lo, 80.0, limit=400, epsrel=1e-10)[0]
        # Gauss points per knot interval integrate the piecewise-linear tables to ~3e-3
        assert abs(q["rates"][0, m] / k - 1) < 5e-3, (pr.type, q["rates"][0, m], k)
