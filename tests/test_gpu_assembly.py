"""GPU parity of the Galerkin assembly kernels against the FP64 oracle (same knots, same tables)."""
import numpy as np
import pytest
import torch

from helpers import oracle_batched, product_grid, product_xsecs
from oracle import operators as op

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _close(a_dev, b, tol=TOL):
    a = a_dev.cpu().numpy()
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max() / scale
    assert err < tol, err
    return err


@pytest.fixture(scope="module")
def grids():
    print("case: assembly Nr=64 lmax=2 (batched grid, snapped knots)", flush=True)
    ops, g, procs = oracle_batched(64, 2)
    pg = product_grid(g)
    return ops, g, procs, pg, product_xsecs()


def test_mass_and_inverse(grids):
    ops, g, procs, pg, xs = grids
    _close(pg.mass_matrix(), ops.M)
    _close(pg.inverse_mass_matrix(), ops.Minv, 1e-10)


def test_advection(grids):
    ops, g, procs, pg, xs = grids
    _close(pg.advection_matrix(), ops.adv_raw)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_collision_operators(grids, k):
    ops, g, procs, pg, xs = grids
    tgK = 300.0 if procs[k].type == "ELASTIC" else 0.0
    ref = op.collision_matrix(g, procs[k], tgK=tgK)
    _close(pg.collision_matrix(xs[k], tgK=tgK), ref)


def test_gas_temperature(grids):
    ops, g, procs, pg, xs = grids
    _close(pg.gas_temperature_matrix(xs[0]), ops.FOp_g)


def test_qoi_vectors(grids):
    ops, g, procs, pg, xs = grids
    _close(pg.mass_op(), ops.mass)
    _close(pg.temp_op(), op.temp_op(g))
    _close(pg.mobility_op(), ops.mobility)
    _close(pg.diffusion_op(xs), ops.diffusion)
    for k in range(3):
        _close(pg.rate_op(xs[k]), ops.rates[k])


def test_maxwellian_init(grids):
    ops, g, procs, pg, xs = grids
    _close(pg.maxwellian_init(), op.maxwellian_init(g, ops.Minv), 1e-11)


def test_basis_table(grids):
    ops, g, procs, pg, xs = grids
    from boltzmann_amd import _lib
    x = np.linspace(0.0, g.t[-1], 777)
    xd = torch.from_numpy(x).cuda()
    for d in (0, 1):
        out = torch.empty((g.num_p, len(x)), dtype=torch.float64, device="cuda")
        _lib.call("bz_basis_table", pg.t, g.num_p, xd, len(x), d, out)
        _close(out, g.V(x, d), 1e-12)
