"""GPU parity of the 1D3V v-space pieces (boltzmann_amd/glow1d.py) against oracle/glow1d.py:
ordinate <-> harmonic transforms and the x-advection half step at the glow-case shapes
(Nr = 128 -> 129 radial B-splines, Nvt = 32, Nx = 200 Chebyshev-Gauss-Lobatto points).
The product computes its own adv_x and eigenvectors from GPU radial integrals, so the
two sides agree to rounding of the eigen decomposition (1e-10 relative)."""
import math

import numpy as np
import pytest
import torch

from oracle import glow1d as gl
from oracle import operators as op

pytestmark = pytest.mark.gpu


def _d(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.mark.parametrize("lmax,Nvt", [(2, 32), (4, 16)])
def test_ords_sph_transforms(lmax, Nvt):
    from boltzmann_amd.glow1d import Ordinates
    print(f"case: ords<->sph lmax={lmax} Nvt={Nvt}", flush=True)
    num_p, Nx = 129, 200
    o = Ordinates(num_p, lmax, Nvt, torch.device("cuda"))
    th, w = gl.ordinates(Nvt)
    po2sh, psh2o = gl.sph_ops(lmax, th, w)
    rng = np.random.default_rng(lmax)
    v = rng.normal(size=(num_p * Nvt, Nx))
    s = rng.normal(size=(num_p * (lmax + 1), Nx))
    a = o.ords_to_sph(_d(v)).cpu().numpy()
    b = o.sph_to_ords(_d(s)).cpu().numpy()
    ra = gl.ords_to_sph(po2sh, v, num_p, Nvt, Nx)
    rb = gl.sph_to_ords(psh2o, s, num_p, lmax + 1, Nx)
    assert np.abs(a - ra).max() <= 1e-13 * np.abs(ra).max()
    assert np.abs(b - rb).max() <= 1e-13 * np.abs(rb).max()


def test_xstep_glow_shapes():
    from boltzmann_amd.glow1d import XAdvection
    from boltzmann_amd.vspace import VGrid
    print("case: xstep Nr=128 Nvt=32 Nx=200", flush=True)
    g = op.grid_single(128, 0, 4.0, 80.0, 5)
    Nvt, Nx, dt, scale = 32, 200, 5e-5, 40.0
    xg = -np.cos(np.pi * np.arange(Nx) / (Nx - 1))
    vg = VGrid(g.t, g.num_p, 0, 5, g.vth, device=torch.device("cuda"))
    xa = XAdvection(vg, xg, Nvt, scale, dt)
    th, _ = gl.ordinates(Nvt)
    xs = gl.XStep(gl.adv_x_operator(g, scale), xg, th, dt)
    rng = np.random.default_rng(5)
    v = rng.normal(size=(g.num_p * Nvt, Nx))
    got = xa.step(_d(v)).cpu().numpy()
    ref = xs.step(v)
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()


def test_strang_steps_sinusoidal_field():
    """The glow configuration's velocity-space path over 20 Strang steps (x half step, v step,
    x half step) at its shapes: Nr = 128, lmax = 2, spline_qpts = 11, Nvt = 32, Nx = 200
    Chebyshev-Gauss-Lobatto points on the 2.54 cm gap, dt = 2.5e-3 RF periods at 13.56 MHz.
    The field is sinusoidal in time with the 100 V / 2.54 cm amplitude, times a cubic
    spatial profile 1 + 0.5 x^3 so the GMRES branch of the v-step runs.  Against the oracle
    with its own operators (oracle/glow1d x-step, dense per-point v-step solves): the
    ordinate state must agree to 1e-8 after every step."""
    from boltzmann_amd.glow1d import Ordinates, VStep, XAdvection
    from boltzmann_amd.solver0d import Batched0D
    from helpers import oracle_batched, product_grid, product_xsecs
    from oracle import krylov as kr
    print("case: 20 Strang steps Nr=128 lmax=2 q=11 Nvt=32 Nx=200, sinusoidal field", flush=True)
    ops, g, procs = oracle_batched(128, 2, q=11)
    s = Batched0D(product_grid(g), product_xsecs())
    Minv = torch.linalg.inv(s.M)
    tau, np0, Tg, dt = 1.0 / 13.56e6, 8e16, 300.0, 2.5e-3
    n0 = 3.22e22 / np0
    Nx, Nvt, L = 200, 32, 2.54e-2
    xg = -np.cos(np.pi * np.arange(Nx) / (Nx - 1))
    ords = Ordinates(g.num_p, 2, Nvt, torch.device("cuda"))
    vs = VStep(Minv @ sum(s.C), Minv @ s.Cg, Minv @ s.A, tau, n0, np0, Tg, dt, ordinates=ords, rtol=1e-12)
    scale = g.vth * tau / 0.5 / L
    xa = XAdvection(s.g, xg, Nvt, scale, 0.5 * dt)
    # oracle
    th, w = gl.ordinates(Nvt)
    po2sh, psh2o = gl.sph_ops(2, th, w)
    xs = gl.XStep(gl.adv_x_operator(g, scale), xg, th, 0.5 * dt)
    K0 = 3.22e22 * (ops.c_en.sum(axis=0) + Tg * ops.c_gT)
    v0 = gl.sph_to_ords(psh2o, ops.initial(Nx), g.num_p, g.num_sh, Nx)
    v_ref, v = v0.copy(), _d(v0)
    for k in range(20):
        E = (100.0 / L) * math.sin(2.0 * math.pi * (k + 0.5) * dt) * (1.0 + 0.5 * xg ** 3)
        v = xa.step(v, 0.5 * dt)
        v = vs.step(v.contiguous(), _d(E))
        v = xa.step(v, 0.5 * dt)
        v_ref = xs.step(v_ref)
        u = gl.ords_to_sph(po2sh, v_ref, g.num_p, Nvt, Nx)
        u = kr.vstep_exact(K0, ops.adv, E, dt, tau, u)
        v_ref = gl.sph_to_ords(psh2o, u, g.num_p, g.num_sh, Nx)
        v_ref = xs.step(v_ref)
        err = np.abs(v.cpu().numpy() - v_ref).max() / np.abs(v_ref).max()
        print(f"step {k}: gmres iterations {vs.last_iters}, max rel diff {err:.2e}", flush=True)
        assert err < 1e-8, (k, err)


def test_glow1d_vspace_dropin_methods():
    """The reference-signature methods (initialize_bte_adv_x, step_bte_x, step_bte_v with
    bs_E set by the caller) drive the same x-v-x step as the classes underneath: three Strang
    steps at the glow shapes against XAdvection + VStep called directly."""
    import argparse
    from boltzmann_amd.glow1d import Ordinates, VStep, XAdvection, glow1d_vspace
    from boltzmann_amd.solver0d import Batched0D
    from helpers import oracle_batched, product_grid, product_xsecs
    print("case: glow1d_vspace drop-in methods, 3 Strang steps", flush=True)
    ops, g, procs = oracle_batched(128, 2, q=11)
    s = Batched0D(product_grid(g), product_xsecs())
    Minv = torch.linalg.inv(s.M)
    C_en, C_gT, A_v = Minv @ sum(s.C), Minv @ s.Cg, Minv @ s.A
    tau, np0, Tg, dt, L = 1.0 / 13.56e6, 8e16, 300.0, 5e-5, 2.54e-2
    Nx, Nvt = 200, 32
    xg = -np.cos(np.pi * np.arange(Nx) / (Nx - 1))
    param = argparse.Namespace(tau=tau, n0=3.22e22 / np0, np0=np0, Tg=Tg, L=L)
    args = argparse.Namespace(rtol=1e-12, atol=1e-20, gmres_rsrt=30)
    gv = glow1d_vspace(C_en, C_gT, A_v, s.g, xg, Nvt, param, args)
    gv.initialize_bte_adv_x(0.5 * dt)
    ords = Ordinates(g.num_p, 2, Nvt, torch.device("cuda"))
    vs = VStep(C_en, C_gT, A_v, tau, param.n0, np0, Tg, dt, ordinates=ords, rtol=1e-12)
    xa = XAdvection(s.g, xg, Nvt, g.vth * tau / (0.5 * L), 0.5 * dt)
    v = ords.sph_to_ords(s.initial(Nx).contiguous())
    w = v.clone()
    for k in range(3):
        E = _d(1e4 * xg ** 3 * (1.0 + 0.1 * k))
        gv.bs_E = E
        v = gv.step_bte_x(v, k * dt, 0.5 * dt)
        v = gv.step_bte_v(v, None, k * dt, dt, "BE")
        v = gv.step_bte_x(v, k * dt, 0.5 * dt)
        w = xa.step(vs.step(xa.step(w, 0.5 * dt).contiguous(), E), 0.5 * dt)
        err = float((v - w).abs().max() / w.abs().max())
        assert err < 1e-12, (k, err)


@pytest.mark.parametrize("m,k,n", [(129, 129, 6400), (17, 5, 33), (1, 1, 1), (300, 77, 50), (16, 4, 16)])
def test_dgemm_nn_matches_fp64_product(m, k, n):
    """bz_dgemm_nn (FP64 MFMA tiles, clamped edges) against the FP64 product on the host."""
    import numpy as np
    import torch
    from boltzmann_amd import _lib
    print(f"case: dgemm m={m} k={k} n={n}", flush=True)
    rng = np.random.default_rng(m * 1000 + k * 10 + n)
    A = rng.normal(size=(m, k))
    B = rng.normal(size=(k, n))
    C = torch.empty((m, n), dtype=torch.float64, device="cuda")
    _lib.call("bz_dgemm_nn", torch.from_numpy(A).cuda(), m, k, torch.from_numpy(B).cuda(), n, C)
    want = A @ B
    scale = np.abs(A) @ np.abs(B)
    err = np.abs(C.cpu().numpy() - want) / np.maximum(scale, 1e-300)
    assert err.max() < 4 * k * 2.2e-16, err.max()
