"""The coupled glow-discharge step (boltzmann_amd.glowdischarge_boltzmann_1d) against the oracle
(oracle/glow_coupled.py): the ion-fluid launch (Poisson fields, backward-Euler step with
the Newton iteration of the reference) at every register layout of the kernel, the
moments that feed it, Strang steps of the whole coupled system, and the time loop with
its checkpoints and restart."""
import math

import numpy as np
import pytest
import torch

from conftest import XSEC_FILE
from helpers import EV_MAX, TE, oracle_batched
from oracle import glow_coupled as gc

pytestmark = pytest.mark.gpu


def _glow(Nr, lmax, Np, Nvt, dt, q=5, **kw):
    from boltzmann_amd.glowdischarge_boltzmann_1d import default_args, glow1d_boltzmann
    opts = dict(collisions=XSEC_FILE, l_max=lmax, Nr=Nr, Np=Np, Nvt=Nvt, Te=TE, ev_max=EV_MAX, ev_extend=1,
                spline_qpts=q, cfl=dt, rtol=1e-12, atol=1e-30, gmres_rsrt=30, max_iter=50)
    opts.update(kw)
    return glow1d_boltzmann(default_args(**opts))


def _oracle(glow, Nr, lmax, q, dt):
    ops, g, procs = oracle_batched(Nr, lmax, q=q)
    np.testing.assert_array_equal(glow.vgrid.t_np, g.t)
    ion = [i for i, p in enumerate(procs) if p.type == "IONIZATION"][0]
    a = glow.args
    return gc.GlowCoupled(ops, g, procs, glow.Np, glow.Nvt, dt, gc.Parameters(), ion, a.atol, a.rtol, a.max_iter)


def _rel(a, b):
    a = a.cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("Np", [20, 50, 100, 137, 200])
def test_fluid_launch_against_oracle(Np):
    """One launch with STEP | FIELD_AFTER on smooth perturbed states: the frozen field, the
    new ion density (oracle with the product's Chebyshev tables: the same arithmetic up to
    rounding; and with the reference's Vandermonde tables) and the v-step field.  Two cases:
    the glow's half step (h = 5e-4, densities of the reference's initial state, J close to
    I) and a long step at 300x the density (h = 0.1, |J - I| up to 25, cond(J) ~ 50)."""
    from boltzmann_amd.glowdischarge_boltzmann_1d import GLOW_FIELD_AFTER, GLOW_STEP
    glow = _glow(8, 1, Np, 4, 1e-3)
    glow._setup_fluid()
    p = glow.param
    xp = glow.xp
    s = 0.5 * (xp + 1)
    ref = gc.cheb_collocation_1d(Np)
    mu, Di = np.full(Np, p.mu_i), np.full(Np, p.Di)
    # a third step length per tile-layout size makes J dominant with a column ratio between 0.5
    # and 1 (host scan of the same J): the Gauss-Jordan tile path instead of the Jacobi one
    h_mid = {100: 0.02, 137: 0.01, 200: 0.002}.get(Np)
    cases = ((5e-4, 1.0), (0.1, 300.0)) + (((h_mid, 1.0),) if h_mid else ())
    for h, amp in cases:
        print(f"case: fluid launch Np={Np} h={h} density x{amp}", flush=True)
        ne = amp * 1e6 * (1e7 + 1e9 * s ** 2 * (1 - s) ** 2) / p.np0 * (1 + 0.02 * np.cos(5 * xp))
        ni = ne * (1 + 0.03 * np.sin(3 * np.pi * xp))
        ki = 2e-3 * (1.5 + 0.5 * np.cos(2 * xp))
        u = np.stack([ne, ni, np.full(Np, 3.0)], axis=1)
        ud = torch.from_numpy(u.copy()).cuda()
        glow._ki.copy_(torch.from_numpy(ki))
        t0, t1 = 0.137, 0.137 + h
        glow._fluid_launch(ud, GLOW_STEP | GLOW_FIELD_AFTER, h=h, t_before=t0, t_after=t1)
        torch.cuda.synchronize()
        st = glow._status.tolist()
        path = "pivoted inverse" if st[3] & 2 else ("Jacobi" if st[3] & 4 else "pivot-free inverse")
        print(f"  status {st} ({path})", flush=True)
        assert st[0] == 1 and (st[3] & 1) == 0, st
        if h < 1e-3 and Np > 90:
            assert (st[3] & 6) == 4, "the glow half step must take the Jacobi tile path"
        if h == h_mid:
            assert (st[3] & 6) == 0, "a column ratio in (0.5, 1) must take the Gauss-Jordan tile path"
        if h > 5e-2 and Np == 200:
            assert (st[3] & 2) != 0, "the long step must take the pivoted path"

        # the field is the derivative of a nearly linear potential: its rounding is bounded by
        # eps |Dp| |LpD^-1| |r| / |E| (kappa ~ 1e4-1e5 here), for the kernel's precomputed
        # -Dp LpD^-1 as for the reference's two products; the two table constructions differ
        # by ~1e-12 of |Dp|, amplified by the same factor
        r = -p.alpha * (ni - ne)
        r[0], r[-1] = math.sin(2 * math.pi * t0), 0.0
        kappa = float((np.abs(glow.Dp) @ (np.abs(glow.LpD_inv) @ np.abs(r))).max() / np.abs(glow._Ef.cpu().numpy()).max())
        for Dp, LpD_inv, e_tab in ((glow.Dp, glow.LpD_inv, 2.2e-16), (ref[1], ref[4], 1e-12)):
            tol_E = 20 * e_tab * kappa
            E = Dp @ (-gc.solve_poisson(LpD_inv, p.alpha, ne, ni, t0))
            print(f"  tables eps {e_tab:g}: kappa {kappa:.2e} field {_rel(glow._Ef, E):.2e} (bound {tol_E:.2e})",
                  flush=True)
            assert _rel(glow._Ef, E) < tol_E
            u1, info = gc.step_fluid_be(Dp, mu, Di, p.n0, ki, E, u, h, glow.args.atol, glow.args.rtol,
                                        glow.args.max_iter)
            dn = _rel(ud[:, 1] - torch.from_numpy(ni).cuda(), u1[:, 1] - ni)
            print(f"  ni {_rel(ud[:, 1], u1[:, 1]):.2e} increment {dn:.2e}", flush=True)
            assert _rel(ud[:, 1], u1[:, 1]) < max(1e-13, tol_E)
            assert dn < max(1e-11, 10 * tol_E)
            bsE = -(Dp @ gc.solve_poisson(LpD_inv, p.alpha, ne, u1[:, 1], t1)) * p.V0 / p.L
            assert _rel(glow.bs_E, bsE) < tol_E
        np.testing.assert_array_equal(ud[:, 0].cpu().numpy(), ne)
        np.testing.assert_array_equal(ud[:, 2].cpu().numpy(), 3.0)
        # without the tile-ordered tables (a C caller may pass none): the same bits
        tiles = glow._fl.DpTile, glow._fl.DppTile
        glow._fl.DpTile = glow._fl.DppTile = None
        u2 = torch.from_numpy(u.copy()).cuda()
        glow._fluid_launch(u2, GLOW_STEP | GLOW_FIELD_AFTER, h=h, t_before=t0, t_after=t1)
        torch.cuda.synchronize()
        glow._fl.DpTile, glow._fl.DppTile = tiles
        assert torch.equal(u2, ud)


def test_fluid_failure_is_reported():
    """A tolerance the iteration cannot reach (rtol 1e-30, 1 iteration per damping level):
    every damping level fails, the launch leaves ni unchanged and flags the failure, and the
    host raises as the reference exits."""
    from boltzmann_amd.glowdischarge_boltzmann_1d import GLOW_STEP
    print("case: fluid failure path Np=50", flush=True)
    Np = 50
    glow = _glow(8, 1, Np, 4, 1e-3)
    glow._setup_fluid()
    glow._fl.rtol, glow._fl.atol, glow._fl.max_iter = 1e-30, 0.0, 1
    s = 0.5 * (glow.xp + 1)
    ne = 1e6 * (1e7 + 1e9 * s ** 2 * (1 - s) ** 2) / glow.param.np0
    u = torch.from_numpy(np.stack([ne, 1.01 * ne, np.full(Np, 3.0)], axis=1)).cuda()
    u0 = u.clone()
    glow._ki.fill_(1e-3)
    glow._fluid_launch(u, GLOW_STEP, h=5e-4, t_before=0.25)
    torch.cuda.synchronize()
    assert int(glow._status[0]) == 0 and int(glow._status[2]) == 1
    assert glow._info[2] <= 1e-6
    assert torch.equal(u, u0)
    with pytest.raises(RuntimeError, match="FAILED"):
        glow.check_fluid_status()
    glow.check_fluid_status()            # the failure count was reset


def test_moments_against_oracle():
    """bte_to_fluid on a positive, non-Maxwellian ordinate state: ne, Te and the ionisation
    rate at every point."""
    print("case: moments Nr=32 lmax=2 Np=50 Nvt=16", flush=True)
    Nr, lmax, q, Np, Nvt = 32, 2, 5, 50, 16
    glow = _glow(Nr, lmax, Np, Nvt, 1e-3, q=q)
    glow._setup_fluid()
    G = _oracle(glow, Nr, lmax, q, 1e-3)
    u0, v0 = G.initialize(TE)
    rng = np.random.default_rng(4)
    v = v0 * (1.0 + 0.2 * rng.uniform(size=v0.shape))
    ud = torch.from_numpy(u0.copy()).cuda()
    glow.bte_to_fluid(ud, torch.from_numpy(v).cuda(), 0.0, 1e-3)
    torch.cuda.synchronize()
    u_ref = G.bte_to_fluid(u0, v)
    assert _rel(ud[:, 0], u_ref[:, 0]) < 1e-13
    assert _rel(ud[:, 2], u_ref[:, 2]) < 1e-13
    assert _rel(glow._ki, G.ki) < 1e-13


@pytest.mark.parametrize("Nr,lmax,q,Np,Nvt,steps", [(32, 2, 5, 50, 16, 8), (128, 2, 11, 200, 32, 3)])
def test_strang_steps_against_oracle(Nr, lmax, q, Np, Nvt, steps):
    """Coupled Strang steps from the reference's initial state (dt = 1e-3 RF periods at
    13.56 MHz, 100 V): u = (ne, ni, Te) and the ordinate state after every step.  The
    v-step GMRES runs to 1e-12; the oracle solves each point exactly."""
    print(f"case: coupled Strang steps Nr={Nr} lmax={lmax} q={q} Np={Np} Nvt={Nvt}", flush=True)
    dt = 1e-3
    glow = _glow(Nr, lmax, Np, Nvt, dt, q=q)
    G = _oracle(glow, Nr, lmax, q, dt)
    u, v = glow.initialize()
    uo, vo = G.initialize(TE)
    assert _rel(u, uo) < 1e-14 and _rel(v, vo) < 1e-13
    u, v = glow.step_init(u, v, dt)
    uo = G.step_init(uo, vo)
    # at t = 0 ni = ne and phi(-1) = 0: the field is rounding (~1e-12 V/m) on both sides
    assert _rel(u, uo) < 1e-12
    assert np.abs(glow.bs_E.cpu().numpy() - G.bs_E).max() < 1e-12 * glow.param.V0 / glow.param.L
    t = 0.0
    for k in range(steps):
        u, v = glow.step(u, v, None, None, t, dt)
        uo, vo = G.step(uo, vo, t)
        t += dt
        eu = [_rel(u[:, c], uo[:, c]) for c in range(3)]
        ev = _rel(v, vo)
        print(f"step {k}: ne {eu[0]:.2e} ni {eu[1]:.2e} Te {eu[2]:.2e} v {ev:.2e} gmres {glow._vstep.last_iters}",
              flush=True)
        assert max(eu) < 1e-9 and ev < 1e-8, (k, eu, ev)


def test_solve_loop_checkpoint_restart(tmp_path):
    """solve over 4 steps writes the reference's checkpoints every 2 steps; restoring the
    second one and solving on reproduces the uninterrupted run bit for bit."""
    print("case: solve loop + restart Nr=16 lmax=1 Np=24 Nvt=8", flush=True)
    dt = 1e-3
    kw = dict(cycles=4 * dt, io_cycle_freq=2 * dt, cp_cycle_freq=2 * dt, dir=str(tmp_path), fname="run")
    glow = _glow(16, 1, 24, 8, dt, **kw)
    u0, v0 = glow.initialize()
    u, v = glow.solve(u0, v0, output_cycle_averaged_qois=True)
    for idx in (0, 1):
        assert (tmp_path / f"run_{idx:04d}_u.npy").exists() and (tmp_path / f"run_{idx:04d}_v.npy").exists()
    g2 = _glow(16, 1, 24, 8, dt, restore=1, rs_idx=1, **kw)
    u1, v1 = g2.initialize()
    u2, v2 = g2.solve(u1, v1)
    assert torch.equal(u2, u) and torch.equal(v2, v)
    # the npz checkpoint of bte_1d3v_solver
    glow.store_checkpoint(v, 4 * dt, dt, str(tmp_path / "cp"))
    tt, dd, vv = glow.restore_checkpoint(str(tmp_path / "cp"))
    assert tt == 4 * dt and dd == dt and torch.equal(vv, v)
    with np.load(tmp_path / "cp.npz") as F:
        ne = F["ne[m^-3]"] / glow.param.np0
    np.testing.assert_allclose(ne, u[:, 0].cpu().numpy(), rtol=1e-12)
