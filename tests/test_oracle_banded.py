"""Pins of oracle/banded.py (the 1D3V batched sweeps of rawkernel.py and the BSR
upper solve of spsm_bsr_upper_tri_solve.cu) against dense linear algebra:
scipy.linalg.solve_triangular, dense products and dense block solves."""
import numpy as np
import pytest
import scipy.linalg

from oracle import banded as bd


def dense_from_diags(diags_s, lower):
    nd, N = diags_s.shape
    A = np.zeros((N, N))
    for d in range(nd):
        for i in range(N):
            j = i - d if lower else i + d
            if 0 <= j < N:
                A[i, j] = diags_s[d, i]
    return A


def rand_diags(rng, bsz, nd, N):
    d = rng.uniform(-1, 1, (bsz, nd, N))
    d[:, 0, :] = rng.uniform(2.0, 3.0, (bsz, N)) * rng.choice([-1, 1], (bsz, N))
    return d


@pytest.mark.parametrize("lower", [True, False])
@pytest.mark.parametrize("nd", [1, 2, 3, 5])
def test_banded_solve_and_mv_against_dense(lower, nd):
    rng = np.random.default_rng(nd + 10 * lower)
    bsz, N = 7, 23
    diags = rand_diags(rng, bsz, nd, N)
    b = rng.normal(size=(bsz, N))
    x = bd.banded_tri_solve(diags, b, lower)
    y = bd.banded_tri_mv(diags, b, lower)
    for s in range(bsz):
        A = dense_from_diags(diags[s], lower)
        ref = scipy.linalg.solve_triangular(A, b[s], lower=lower)
        np.testing.assert_allclose(x[s], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        np.testing.assert_allclose(y[s], A @ b[s], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("lower", [True, False])
def test_bidiagonal_against_dense(lower):
    rng = np.random.default_rng(4)
    bsz, N = 5, 31
    diag = rng.uniform(1.5, 2.5, (bsz, N))
    sdiag = rng.uniform(-1, 1, (bsz, N - 1))
    b = rng.normal(size=(bsz, N))
    x = bd.bidiagonal_solve(diag, sdiag, b, lower)
    for s in range(bsz):
        A = np.diag(diag[s]) + (np.diag(sdiag[s], -1) if lower else np.diag(sdiag[s], 1))
        np.testing.assert_allclose(A @ x[s], b[s], rtol=1e-12, atol=1e-12)


def adv_dense(D, av, dt, lower):
    nd, N = D.shape
    A = np.eye(N)
    for i in range(N):
        for d in range(nd):
            j = i - d if lower else i + d
            if 0 <= j < N:
                A[i, j] += dt * av * (D[d, j] if d > 0 else D[0, i])
    return A


@pytest.mark.parametrize("kind", ["vr", "vt", "x"])
def test_advection_sweeps_against_dense(kind):
    rng = np.random.default_rng({"vr": 1, "vt": 2, "x": 3}[kind])
    Nr, Nvt, Nx, nd, dt = 9, 6, 5, 3, 0.3
    Ex = rng.normal(size=Nx) * 2
    cos_vt = np.cos(np.linspace(0.1, np.pi - 0.1, Nvt))
    vr = rng.uniform(0.5, 2.0, Nr)
    N = {"vr": Nr, "vt": Nvt, "x": Nx}[kind]
    DL = rng.uniform(0, 1, (nd, N))
    DR = rng.uniform(0, 1, (nd, N))
    if kind == "vr":
        b = rng.normal(size=(Nx * Nvt, Nr))
        out = bd.adv_vr_BE(DL, DR, Ex, cos_vt, dt, b, Nr, Nvt, Nx)
        av = (-Ex[:, None] * cos_vt[None, :]).reshape(-1)
        fwd = av >= 0
    elif kind == "vt":
        b = rng.normal(size=(Nx * Nr, Nvt))
        out = bd.adv_vt_BE(DL, DR, Ex, vr, dt, b, Nr, Nvt, Nx)
        av = (Ex[:, None] / vr[None, :]).reshape(-1)
        fwd = -av >= 0
    else:
        b = rng.normal(size=(Nr * Nvt, Nx))
        out = bd.adv_x_BE(DL, DR, vr, cos_vt, dt, b, Nr, Nvt, Nx)
        av = (vr[:, None] * cos_vt[None, :]).reshape(-1)
        fwd = av >= 0
    assert fwd.any() and (~fwd).any()
    for s in range(b.shape[0]):
        A = adv_dense(DL if fwd[s] else DR, av[s], dt, fwd[s])
        np.testing.assert_allclose(A @ out[s], b[s], rtol=1e-11, atol=1e-11 * np.abs(b[s]).max())


def test_bsr_upper_unit_solve_against_dense():
    rng = np.random.default_rng(8)
    nbr, bsz, nrhs = 6, 4, 3
    blocks = {}
    for i in range(nbr):
        for j in range(max(0, i - 1), min(nbr, i + 3)):
            blocks[(i, j)] = rng.normal(size=(bsz, bsz)) * 0.3
    indptr, indices, data = [0], [], []
    for i in range(nbr):
        for j in sorted(jj for (ii, jj) in blocks if ii == i):
            indices.append(j)
            data.append(blocks[(i, j)])
        indptr.append(len(indices))
    data = np.array(data)
    U = np.eye(nbr * bsz)
    for (i, j), B in blocks.items():
        if j > i:
            U[i * bsz:(i + 1) * bsz, j * bsz:(j + 1) * bsz] = B
    b = rng.normal(size=(nbr * bsz, nrhs))
    x = bd.bsr_upper_unit_solve(data, np.array(indptr), np.array(indices), b)
    np.testing.assert_allclose(U @ x, b, rtol=1e-12, atol=1e-12)
