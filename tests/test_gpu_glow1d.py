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
