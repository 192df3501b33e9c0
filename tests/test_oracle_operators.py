"""Pins of the oracle's Galerkin operators and QoI vectors.

References restated by oracle/operators.py: spec_spherical.py:226-336, 442-498,
advection_operator_spherical_polys.py:6-16, collision_operator_spherical.py:96-225,
443-497, utils.py:73-150, 321-395, 475-532, 854-944.  Checked here against
  * exact rational integrals (sympy, tests/golden/bspline_exact.npz) of the mass,
    advection, constant-cross-section loss and gas-temperature radial blocks;
  * the spherical-harmonic coupling integrals <Y_q, mu Y_l> and
    <Y_q, (1 - mu^2) dY_l/dmu> computed by Gauss-Legendre quadrature in mu;
  * the partition of unity: M^-1 (mass / sqrt(4 pi)) = 1 on l = 0; the l = 0 collision
    block conserves particles (elastic, excitation) or creates one electron per
    ionization (column sum = ionization rate integral);
  * the energy moment: with v^2 = sum_p c_p B_p, c^T q0 gives the energy lost per
    collision (-threshold for excitation/ionization, -(2m/M) v^2 for elastic), which
    fixes the post-collision speed v' of every process;
  * closed forms of the QoIs of a Maxwellian (mean energy 3/2 Te, rate sigma0 <v>,
    diffusion <v>/(3 sigma0) for constant cross sections).
"""
import math
import os

import numpy as np
import pytest

from helpers import oracle_batched, oracle_procs
from oracle import operators as op
from oracle import solver0d as s0

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "bspline_exact.npz"))


def golden_grid(name, lmax=2, q=5, vth=2.0):
    t = GOLD[f"{name}_knots"]
    return op.Grid(t, len(t) - 4, lmax, q, vth)


def const_process(ptype, sigma0, threshold=0.0):
    return op.Process(ptype, [0.0, 1e6], [sigma0, sigma0], threshold=threshold, mass_ratio=1.36e-5)


@pytest.mark.parametrize("name", ["uniform", "snapped"])
def test_mass_and_advection_blocks_exact(name):
    g = golden_grid(name)
    M = op.mass_matrix(g)
    ns = g.num_sh
    for l in range(ns):
        np.testing.assert_allclose(M[l::ns, l::ns], GOLD[f"{name}_M"], rtol=0, atol=1e-14)
        for m in range(ns):
            if m != l:
                assert (M[l::ns, m::ns] == 0).all()
    mm1, mm2 = op.advection_radial(g)
    np.testing.assert_allclose(mm1, GOLD[f"{name}_mm1"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(mm2, GOLD[f"{name}_mm2"], rtol=0, atol=1e-14)


def sh_coupling(lmax):
    """a[q,l] = <Y_q, mu Y_l>, d[q,l] = <Y_q, (1-mu^2) dY_l/dmu> (m = 0, orthonormal)."""
    mu, w = np.polynomial.legendre.leggauss(40)
    Y = []
    dY = []
    for l in range(lmax + 1):
        c = np.zeros(l + 1)
        c[l] = 1.0
        nl = math.sqrt((2 * l + 1) / (4 * math.pi))
        Y.append(nl * np.polynomial.legendre.legval(mu, c))
        dY.append(nl * np.polynomial.legendre.legval(mu, np.polynomial.legendre.legder(c)) if l else 0 * mu)
    a = np.array([[2 * math.pi * np.dot(w, Y[q] * mu * Y[l]) for l in range(lmax + 1)] for q in range(lmax + 1)])
    d = np.array([[2 * math.pi * np.dot(w, Y[q] * (1 - mu ** 2) * dY[l]) for l in range(lmax + 1)]
                  for q in range(lmax + 1)])
    return a, d


@pytest.mark.parametrize("lmax", [1, 2, 4])
def test_advection_coupling_is_the_angular_projection(lmax):
    g = golden_grid("uniform", lmax=lmax)
    A = op.advection_matrix(g)
    mm1, mm2 = GOLD["uniform_mm1"], GOLD["uniform_mm2"]
    a, d = sh_coupling(lmax)
    ns = g.num_sh
    tol = 1e-13 * max(np.abs(mm1).max(), np.abs(mm2).max())
    for qi in range(ns):
        for l in range(ns):
            np.testing.assert_allclose(A[qi::ns, l::ns], a[qi, l] * mm1 + d[qi, l] * mm2, rtol=0, atol=tol)


def test_inverse_mass_and_partition_of_unity():
    ops, g, _ = oracle_batched(32, 2)
    ns = g.num_sh
    e0 = np.zeros(g.n)
    e0[0::ns] = 1.0
    np.testing.assert_allclose(ops.Minv @ (ops.mass / math.sqrt(4 * math.pi)), e0, rtol=0, atol=1e-10)
    np.testing.assert_allclose(ops.Minv @ ops.M, np.eye(g.n), rtol=0, atol=1e-9)


@pytest.mark.parametrize("name", ["uniform", "snapped"])
def test_constant_sigma_loss_and_gas_temperature_exact(name):
    g = golden_grid(name)
    s0_ = 3.3e-20
    q0, q1 = op.collision_radial(g, const_process("ELASTIC", s0_))
    ref = -g.vth * s0_ * GOLD[f"{name}_S3"]
    np.testing.assert_allclose(q1, ref, rtol=0, atol=1e-14 * np.abs(ref).max())
    gt = op.gas_temperature_radial(g, const_process("ELASTIC", s0_))
    D3 = GOLD[f"{name}_D3"]
    ratio = gt[np.abs(D3) > 1e-3 * np.abs(D3).max()] / D3[np.abs(D3) > 1e-3 * np.abs(D3).max()]
    # the shape is -kappa sigma0 int x^3 B_p' B_k' with one positive constant kappa
    assert np.ptp(ratio) <= 1e-12 * np.abs(ratio).max() and ratio[0] < 0


def v2_coefficients(g):
    """c with sum_p c_p B_p(x) = x^2 (x^2 lies in the cubic spline space)."""
    x = np.linspace(g.t[0], g.t[-1], 4 * g.num_p)
    V = g.V(x)
    c, *_ = np.linalg.lstsq(V.T, x ** 2, rcond=None)
    assert np.abs(V.T @ c - x ** 2).max() < 1e-11 * x.max() ** 2
    return c


def test_collision_particle_and_energy_balance():
    procs = oracle_procs()
    g = op.grid_batched(48, 1, 4.0, 80.0, 5, [p.threshold for p in procs])
    c2 = v2_coefficients(g)
    xg, wg = g.quad(g.q)
    ev = (xg * g.vth / op.C_GAMMA) ** 2
    V = g.V(xg)
    for p in procs:
        q0, _ = op.collision_radial(g, p, tgK=0.0)
        sig = p.total_cross_section(ev)
        loss = V @ (wg * g.vth * xg ** 3 * sig)            # int vth x^3 sigma B_k
        colsum = q0.sum(axis=0)
        energy = c2 @ q0
        scale = np.abs(loss).max()
        if p.type == "ELASTIC":
            cmu = 2.0 * p.mByM
            np.testing.assert_allclose(colsum, 0.0, atol=1e-12 * scale)
            np.testing.assert_allclose(energy, -cmu * (V @ (wg * g.vth * xg ** 5 * sig)),
                                       rtol=0, atol=1e-9 * cmu * scale * xg.max() ** 2)
        else:
            keep = ev >= p.threshold
            loss_k = V[:, keep] @ (wg * g.vth * xg ** 3 * sig)[keep]
            xth2 = p.threshold * op.C_GAMMA ** 2 / g.vth ** 2
            gain = 1.0 if p.type == "IONIZATION" else 0.0      # electrons created per collision
            np.testing.assert_allclose(colsum, gain * loss_k, rtol=0, atol=1e-12 * scale)
            np.testing.assert_allclose(energy, -xth2 * loss_k, rtol=0, atol=1e-10 * xth2 * scale)


def test_gas_temperature_zero_column_sum():
    procs = oracle_procs()
    g = op.grid_batched(48, 1, 4.0, 80.0, 5, [p.threshold for p in procs])
    el = [p for p in procs if p.type == "ELASTIC"][0]
    gt = op.gas_temperature_radial(g, el)
    np.testing.assert_allclose(gt.sum(axis=0), 0.0, atol=1e-12 * np.abs(gt).max())


def test_maxwellian_qoi_closed_forms():
    Te = 4.0
    sig0 = 2.0e-20
    g = op.grid_batched(96, 1, Te, 120.0, 5, [], extend=False)
    procs = [const_process("ELASTIC", sig0), const_process("EXCITATION", sig0, threshold=0.0)]
    ops = s0.BatchedOps(g, procs)
    h = ops.initial(1)
    # density of the projected Maxwellian exp(-x^2)/pi^1.5 is 1
    assert ops.mass @ h[:, 0] == pytest.approx(1.0, rel=1e-9)
    prm = dict(n0=np.array([1e22]), Tg=np.array([0.0]), E=np.array([1.0]), ns_by_n0=np.ones((1, 1)))
    hn = s0.normalized_distribution(ops, h)
    q = s0.compute_qois(ops, prm, hn)
    assert q["energy"][0] == pytest.approx(1.5 * Te, rel=1e-9)
    mean_speed = g.vth * 2.0 / math.sqrt(math.pi)
    assert q["rates"][1, 0] == pytest.approx(sig0 * mean_speed, rel=1e-9)
    assert q["diffusion"][0] == pytest.approx(mean_speed / (3.0 * 2.0 * sig0), rel=1e-9)


def _bspline_ld(t, k, x):
    """B_k(x) in long double by the Cox-de Boor recursion on the local knots (half-open
    intervals), for the brute-force check below."""
    LD = np.longdouble
    tk = np.asarray(t[k:k + 5], dtype=LD)
    x = LD(x)
    N = [LD(1) if tk[j] <= x < tk[j + 1] else LD(0) for j in range(4)]
    for d in range(1, 4):
        Nn = []
        for j in range(4 - d):
            a, b = tk[j + d] - tk[j], tk[j + d + 1] - tk[j + 1]
            v = LD(0)
            if a != 0:
                v += (x - tk[j]) / a * N[j]
            if b != 0:
                v += (tk[j + d + 1] - x) / b * N[j + 1]
            Nn.append(v)
        N = Nn
    return N[0]


def test_elastic_gain_difference_without_cancellation():
    """B_p(v sqrt(1 - 2m/M)) - B_p(v), evaluated as the integral of B_p' over [v', v], against a
    long-double evaluation of the two values (relative 1e-12 of the difference, where the
    FP64 difference of the values keeps only ~1e-11), and against the FP64 difference of the
    values to the level of that cancellation."""
    ops, g, procs = oracle_batched(32, 1)
    c_mu = 2.0 * 1.36e-5
    x, _ = g.quad(g.q)
    d = op.elastic_gain_difference(g, x, c_mu)
    naive = g.V(x * math.sqrt(1.0 - c_mu)) - g.V(x)
    scale = np.abs(g.V(x)).max()
    assert np.abs(d - naive).max() <= 1e-9 * scale
    rng = np.random.default_rng(4)
    LD = np.longdouble
    s = np.sqrt(LD(1) - LD(c_mu))
    for j in rng.choice(len(x), 12, replace=False):
        xs = LD(x[j])
        for p in range(g.num_p):
            if not (g.t[p] <= x[j] <= g.t[p + 4]):
                continue
            ref = _bspline_ld(g.t, p, xs * s) - _bspline_ld(g.t, p, xs)
            if ref == 0:
                continue
            assert abs(LD(d[p, j]) - ref) <= LD(1e-12) * abs(ref) + LD(1e-30), (p, j)
