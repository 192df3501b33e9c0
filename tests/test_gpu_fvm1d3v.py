"""The fully implicit 1D3V step on the finite-volume / ordinate grid
(boltzmann_amd.bte_1d3v_solver, kernels bz_fvm_rhs_mvec and bz_fvm_precond_xvtvrc) against the
oracle (oracle/fvm1d3v.py): the host set-up, the step operator, the XVTVRC preconditioner,
the GMRES step against the dense solve, a sequence of steps in a time-varying field with its
moments, and the step's residual at a production size."""
import types

import numpy as np
import pytest
import torch

from conftest import XSEC_FILE
from helpers import oracle_procs
from oracle import fvm1d3v as fv

pytestmark = pytest.mark.gpu


def _table(**kw):
    tp = dict(L=2.54e-2, V0=100.0, freq=13.56e6, Tg=300.0, p0=1.0, collisions=XSEC_FILE, sp_order=3,
              spline_qpts=3, Nr=15, Np=12, Nvt=8, dt=1e-3, Te=4.0, ev_max=80.0, ev_extend=0, atol=1e-12,
              rtol=1e-10, gmres_rsrt=5, gmres_rsrt_iter=10, pvr=3, pvt=3, px=1, solver_type="fully_implicit",
              vspace_type=0, xspace_type=0)
    tp.update(kw)
    return tp


def _pair(**kw):
    from boltzmann_amd.bte_1d3v_solver import bte_1d3v, params
    tp = _table(**kw)
    S = bte_1d3v(types.SimpleNamespace(params=params(tp=tp)))
    keys = ("L", "V0", "freq", "Tg", "p0", "Np", "Nr", "Nvt", "sp_order", "spline_qpts", "ev_max", "ev_extend",
            "Te", "dt", "pvr", "pvt", "px", "atol", "rtol", "gmres_rsrt", "gmres_rsrt_iter")
    O = fv.Fvm1d3v(fv.Params(**{k: tp[k] for k in keys}), oracle_procs())
    return S, O


def _rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / np.abs(b).max())


def _field(x, t=0.25):
    """The reference driver's ef_mode 1: E(x, t) = 1e4 x^3 sin(2 pi t) V/m (both signs)."""
    return 1e4 * x ** 3 * np.sin(2 * np.pi * t)


# the third case has enough radial points (65) for the segmented-scan radial sweep; the last
# two use stencil widths without a fixed-width step-operator kernel (the general one), the
# 1.2x extended domain and two-point radial stencils in the scan
CASES = [dict(), dict(pvr=2, pvt=2, px=2, Nvt=12, ev_extend=2, Np=9), dict(spline_qpts=5, Np=10),
         dict(pvr=1, pvt=2, px=3, Nvt=10, Np=11, ev_extend=1), dict(pvr=1, spline_qpts=5, Np=8)]


@pytest.mark.parametrize("kw", CASES)
def test_setup_matches_oracle(kw):
    S, O = _pair(**kw)
    np.testing.assert_array_equal(S.xp_vt, O.vt)
    assert np.abs(S.xp_vr - O.vr).max() < 1e-15 * O.vr.max()
    assert np.abs(S.G - O.G).max() <= 1e-14 * np.abs(O.G).max()
    assert np.abs(S.loss - O.loss).max() <= 1e-14 * O.loss.max()
    for a, b in ((S.op_mass, O.op_mass), (S.op_temp, O.op_temp), (np.stack(S.op_rate), np.stack(O.op_rate))):
        assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()


@pytest.mark.parametrize("kw", CASES)
def test_rhs_mvec_matches_oracle(kw):
    S, O = _pair(**kw)
    dt = 2e-3
    S.init(dt, 1.0, 0.7)
    O.init(dt, 1.0, 0.7)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(O.dof_v, O.Nx))
    E = _field(O.xp)
    E[O.Nx // 2] = 0.0                       # E = 0 takes the Ep branch
    want = O.rhs_mvec(E, x, dt)
    got = S.rhs_mvec(S._field(E), torch.from_numpy(x.reshape(-1)).cuda(), dt)
    assert _rel(got, want) < 1e-13, _rel(got, want)


# 319 radial points: segments of 10 points, the radial scan's 24-point register form; then
# the angular sweep with 2-, 3- and 5-point stencils and 16 ordinates, and 150 x points
# (the x scan's four positions per lane)
PC_CASES = [dict(Nr=31, spline_qpts=11, Np=6), dict(Nr=31, spline_qpts=11, Np=7, pvr=2, ev_extend=1),
            dict(pvt=1, Nvt=16, Np=9), dict(pvt=2, Np=150), dict(pvt=4, Nvt=12, Np=10)]


@pytest.mark.parametrize("kw", CASES + PC_CASES)
def test_preconditioner_matches_oracle(kw):
    S, O = _pair(**kw)
    dt = 2e-3
    S.init(dt, 0.8, 1.0)
    O.init(dt, 0.8, 1.0)
    rng = np.random.default_rng(4)
    b = rng.normal(size=(O.dof_v, O.Nx))
    for Eamp in (1.0, -3.0):
        E = Eamp * _field(O.xp)
        want = O.step_op_split(E, b, dt, alpha=0.6)
        got = S.step_op_split(S._field(E), torch.from_numpy(b).cuda(), None, 0.0, dt, alpha=0.6)
        assert _rel(got, want) < 1e-12, _rel(got, want)


@pytest.mark.parametrize("device_gmres", [True, False])
@pytest.mark.parametrize("kw", CASES)
def test_step_matches_dense_solve(kw, device_gmres):
    """Both GMRES drivers: the device one (graph-replayed Arnoldi steps, true residual per
    cycle) and the reference's loop (true residual after every step)."""
    S, O = _pair(**kw)
    S.use_device_gmres = device_gmres
    dt = 1e-3
    S.init(dt)
    O.init(dt)
    v = O.initial_condition()
    vd = S.initial_condition()
    assert _rel(vd, v) < 1e-15
    E = _field(O.xp)
    y = S.step(S._field(E), vd, None, 0.0, dt)
    yd = np.linalg.solve(O.rhs_matrix(E, dt), O.rhs_vector(v)).reshape(O.dof_v, O.Nx)
    assert _rel(y, yd) < 1e-8, _rel(y, yd)
    assert S.gmres_info["status"] == 0


def test_time_steps_and_moments():
    """Eight steps in the driver's field E(x, 0.25 + k dt): the product's state and moments
    follow the oracle's GMRES steps."""
    S, O = _pair(Nvt=12)
    dt = 5e-4
    S.init(dt)
    O.init(dt)
    v, vd = O.initial_condition(), S.initial_condition()
    for k in range(8):
        E = _field(O.xp, 0.25 + k * dt)
        v = O.step(E, v, dt)
        vd = S.step(S._field(E), vd, None, k * dt, dt)     # replays the captured GMRES steps
        assert _rel(vd, v) < 1e-8, (k, _rel(vd, v))
    q, qd = O.compute_qoi(v), S.compute_qoi(vd)
    for key in ("ne", "Te"):
        assert _rel(qd[key], q[key]) < 1e-8, key
    assert _rel(qd["rates"], q["rates"]) < 1e-8
    assert _rel(S.normalize_edf(vd), O.normalize_edf(v)) < 1e-8


@pytest.mark.parametrize("device_gmres", [True, False])
def test_gmres_failure_raises(device_gmres):
    S, O = _pair()
    S.use_device_gmres = device_gmres
    dt = 1e-3
    S.params.gmres_maxiter = 1
    S.params.rtol = 1e-30
    S.params.atol = 0.0
    S.init(dt)
    with pytest.raises(RuntimeError, match="GMRES solver failed"):
        S.step(S._field(_field(O.xp)), S.initial_condition(), None, 0.0, dt)


def test_production_size_residual():
    """Nr = 63 splines, 11 points per knot interval (671 radial points), Nvt = 16, Np = 100:
    the step converges and its true residual meets the GMRES tolerance; the step operator is
    linear in the state."""
    from boltzmann_amd.bte_1d3v_solver import bte_1d3v, params
    S = bte_1d3v(types.SimpleNamespace(params=params(tp=_table(Nr=63, spline_qpts=11, Nvt=16, Np=100, ev_extend=2,
                                                               gmres_rsrt=10, gmres_rsrt_iter=20))))
    dt = 2e-4
    S.init(dt)
    v = S.initial_condition()
    E = S._field(_field(S.xp))
    y = S.step(E, v, None, 0.0, dt)
    b = torch.linalg.vector_norm(v)
    assert S.gmres_info["status"] == 0
    assert S.gmres_info["residual_abs"] <= max(S.params.atol, 1.01 * S.params.rtol * float(b))
    x1 = torch.randn(S.dof_v * S.dof_x, dtype=torch.float64, device="cuda")
    x2 = torch.randn_like(x1)
    lhs = S.rhs_mvec(E, 2.0 * x1 - 3.0 * x2, dt)
    rhs = 2.0 * S.rhs_mvec(E, x1, dt) - 3.0 * S.rhs_mvec(E, x2, dt)
    assert float((lhs - rhs).abs().max() / rhs.abs().max()) < 1e-13
    assert torch.isfinite(y).all()
