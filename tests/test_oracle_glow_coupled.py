"""Pins the oracle of the glow-discharge coupling (oracle/glow_coupled.py) on the CPU with
what the mathematics fixes: exact differentiation of polynomials, a manufactured Poisson
solution, the Jacobian of an affine right-hand side, the backward-Euler equation and ion
balance, the damping path of the Newton iteration and the moments of a Maxwellian.  Also
checks the product's host-side Chebyshev tables (closed form) against the oracle's
(Vandermonde, the reference's construction)."""
import math

import numpy as np
import pytest

from helpers import oracle_batched
from oracle import glow_coupled as gc
from oracle import operators as op


def _cheb_poly_values(c, x, m=0):
    """Values of the m-th derivative of the Chebyshev series c at x."""
    P = np.polynomial.Chebyshev(c)
    return P.deriv(m)(x) if m else P(x)


@pytest.mark.parametrize("Np", [8, 33, 200])
def test_collocation_exact_on_polynomials(Np):
    xp, Dp, Lp, LpD, LpD_inv = gc.cheb_collocation_1d(Np)
    assert np.all(np.diff(xp) > 0) and xp[0] == -1.0 and xp[-1] == 1.0
    rng = np.random.default_rng(Np)
    c = rng.normal(size=Np) / (1.0 + np.arange(Np)) ** 2          # degree Np-1
    f = _cheb_poly_values(c, xp)
    for D, m, tol in ((Dp, 1, 1e-11), (Lp, 2, 1e-9)):
        d = _cheb_poly_values(c, xp, m)
        assert np.abs(D @ f - d).max() / np.abs(d).max() < tol, m
    # closed-form differentiation matrix (barycentric weights, negative-sum diagonal)
    x2, D2 = gc.cheb_diff_closed_form(Np)
    np.testing.assert_array_equal(x2, xp)
    assert np.abs(D2 - Dp).max() / np.abs(Dp).max() < 1e-12
    # a planted sign error in the closed form must fail
    bad = D2.copy()
    bad[1, 2] = -bad[1, 2]
    assert np.abs(bad - Dp).max() / np.abs(Dp).max() > 1e-6


@pytest.mark.parametrize("Np", [8, 33, 200])
def test_product_collocation_matches_oracle(Np):
    from boltzmann_amd.collocation_op import cheb_collocation_1d
    xp, Dp, Lp, LpD, LpD_inv = gc.cheb_collocation_1d(Np)
    c = cheb_collocation_1d(Np)
    np.testing.assert_array_equal(c.xp, xp)
    assert np.abs(c.Dp - Dp).max() / np.abs(Dp).max() < 1e-12
    assert np.abs(c.Lp - Lp).max() / np.abs(Lp).max() < 1e-11
    assert np.abs(c.LpD_inv - LpD_inv).max() / np.abs(LpD_inv).max() < 1e-10


def test_poisson_manufactured():
    """phi = sin(2 pi t) (1 - x)/2 + (1 - x^2) q(x): phi'' = -alpha (ni - ne) with ne, ni
    built from phi'' recovers phi at every point."""
    Np = 64
    xp, Dp, Lp, LpD, LpD_inv = gc.cheb_collocation_1d(Np)
    t, alpha = 0.13, 3.7
    bc = math.sin(2 * math.pi * t)
    q = np.polynomial.Chebyshev([0.3, -0.2, 0.05, 0.1])
    x = np.polynomial.Polynomial([0, 1])
    phi = bc * (1 - x) / 2 + (1 - x ** 2) * q.convert(kind=np.polynomial.Polynomial)
    d2 = phi.deriv(2)(xp)
    ne = np.full(Np, 2.0)
    ni = ne - d2 / alpha
    got = gc.solve_poisson(LpD_inv, alpha, ne, ni, t)
    assert np.abs(got - phi(xp)).max() < 1e-12
    assert abs(got[0] - bc) < 1e-13 and abs(got[-1]) < 1e-13


def _fluid_case(Np=48, seed=3):
    rng = np.random.default_rng(seed)
    p = gc.Parameters()
    xp, Dp, Lp, LpD, LpD_inv = gc.cheb_collocation_1d(Np)
    s = 0.5 * (xp + 1)
    ne = 1e3 * (1 + 50 * s ** 2 * (1 - s) ** 2)
    ni = ne * (1 + 0.05 * rng.normal(size=Np))
    ki = 1e-3 * (1 + rng.uniform(size=Np))
    E = 30.0 * np.sin(np.pi * xp) + 5.0 * xp ** 3
    return p, Dp, np.full(Np, p.mu_i), np.full(Np, p.Di), ne, ni, ki, E


def test_fluid_jacobian_is_the_derivative():
    p, Dp, mu, Di, ne, ni, ki, E = _fluid_case()
    J = gc.rhs_fluid_jacobian(Dp, mu, Di, E)
    d = np.random.default_rng(5).normal(size=len(ni)) * ni * 1e-2
    r0 = gc.rhs_fluid(Dp, mu, Di, p.n0, ki, E, ne, ni)
    r1 = gc.rhs_fluid(Dp, mu, Di, p.n0, ki, E, ne, ni + d)
    assert np.abs((r1 - r0) - J @ d).max() / np.abs(J @ d).max() < 1e-10
    # the weak boundary rows: the flux there is mu ni E only
    Jbad = J.copy()
    Jbad[:, 0] *= 1.0 + 1e-6
    assert np.abs((r1 - r0) - Jbad @ d).max() / np.abs(J @ d).max() > 1e-9


def test_fluid_be_step_equation_and_ion_balance():
    """ni1 - ni = h rhs(ni1), and with Clenshaw-Curtis weights w (exact on degree < Np),
    w.(ni1 - ni) = h (w.(ki n0 ne) - (flux(1) - flux(-1)))."""
    p, Dp, mu, Di, ne, ni, ki, E = _fluid_case()
    Np = len(ni)
    u = np.stack([ne, ni, np.full(Np, 3.0)], axis=1)
    h = 5e-4
    u1, info = gc.step_fluid_be(Dp, mu, Di, p.n0, ki, E, u, h, 1e-30, 1e-13, 50)
    assert info["status"] and info["iter"] <= 2
    n1 = u1[:, 1]
    res = n1 - ni - h * gc.rhs_fluid(Dp, mu, Di, p.n0, ki, E, ne, n1)
    assert np.abs(res).max() / np.abs(n1 - ni).max() < 1e-10
    # Clenshaw-Curtis weights on the Lobatto points: integrals of the Lagrange basis
    xp = -np.cos(np.pi * np.arange(Np) / (Np - 1))
    V = np.polynomial.chebyshev.chebvander(xp, Np - 1)
    ints = np.array([(np.polynomial.Chebyshev.basis(k).integ(lbnd=-1))(1.0) for k in range(Np)])
    w = np.linalg.solve(V.T, ints)
    flux = mu * n1 * E - Di * (Dp @ n1)
    flux[0], flux[-1] = mu[0] * n1[0] * E[0], mu[-1] * n1[-1] * E[-1]
    lhs = w @ (n1 - ni)
    rhs = h * (w @ (ki * p.n0 * ne) - (flux[-1] - flux[0]))
    assert abs(lhs - rhs) / abs(h * (flux[-1] - flux[0])) < 1e-9


def test_newton_solver_damping_path():
    """x^3 + x - 10 = 0 from 0 with the Jacobian frozen at 0 (= 1): the undamped and the
    alpha = 1/4 iterations diverge (contraction factors -12 and -2.25 at the root 2), alpha =
    1/16 converges (0.19), so the solver ends at alpha = 0.0625 on the root."""
    res = lambda x: x ** 3 + x - 10.0
    jac = lambda x: np.array([[1.0]])
    info = gc.newton_solver(np.zeros(1), res, jac, 1e-14, 1e-14, 500)
    assert info["status"]
    assert info["alpha"] == 0.0625
    assert abs(info["x"][0] - 2.0) < 1e-12


def test_moments_of_a_maxwellian():
    """bte_to_fluid on ne(x) times a unit-mass Maxwellian at the grid temperature: ne back,
    Te = 4 eV up to the truncation of the v-domain, and ki / (np0 tau) = <sigma v> of the
    Maxwellian, the rate coefficient integrated independently (scipy quad over energy)."""
    import scipy.constants as sc
    import scipy.integrate as si
    ops, g, procs = oracle_batched(32, 1)
    Np, Nvt = 12, 8
    p = gc.Parameters()
    ion = [i for i, pr in enumerate(procs) if pr.type == "IONIZATION"][0]
    G = gc.GlowCoupled(ops, g, procs, Np, Nvt, 1e-3, p, ion)
    u, v = G.initialize(4.0)
    u2 = G.bte_to_fluid(u, v)
    np.testing.assert_allclose(u2[:, 0], u[:, 0], rtol=1e-12)
    assert np.abs(u2[:, 2] - 4.0).max() < 4e-4
    kT = 4.0
    f = lambda e: 2.0 * np.sqrt(e / np.pi) * kT ** -1.5 * np.exp(-e / kT)
    vel = lambda e: np.sqrt(2.0 * e * sc.e / sc.m_e)
    sig = lambda e: float(procs[ion].total_cross_section(np.array([e]))[0])
    th = procs[ion].threshold
    k_phys = si.quad(lambda e: sig(e) * vel(e) * f(e), th, 60.0, limit=400, epsrel=1e-10)[0]
    k_phys += si.quad(lambda e: sig(e) * vel(e) * f(e), 60.0, 80.0, limit=200, epsrel=1e-10)[0]
    got = G.ki / (p.np0 * p.tau)
    assert np.abs(got / k_phys - 1.0).max() < 2e-3, (got[0], k_phys)
    assert np.ptp(got) <= 1e-14 * got[0]
