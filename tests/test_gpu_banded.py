"""GPU parity of the batched 1D3V sweeps (bz_banded.hip) against oracle/banded.py.

Same seeded inputs on both sides; FP64 results agree to 1e-13 relative (the
kernels may contract a*b+c into one fused multiply-add where the numpy oracle
rounds twice).  Sizes cover one and several workgroups, ragged last chunks and
workgroups, single-column systems, and both sweep directions in one launch.
"""
import numpy as np
import pytest
import torch

from oracle import banded as bd

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _d(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _close(got, ref):
    got = got.cpu().numpy() if isinstance(got, torch.Tensor) else got
    scale = np.abs(ref).max() if ref.size else 1.0
    assert np.abs(got - ref).max() <= TOL * max(scale, 1e-300), np.abs(got - ref).max() / scale


@pytest.mark.parametrize("bsz,N", [(1, 2), (5, 17), (64, 40), (130, 200), (300, 33)])
@pytest.mark.parametrize("lower", [True, False])
def test_bidiagonal(bsz, N, lower):
    from boltzmann_amd import rawkernel as rk
    print(f"case: bidiag bsz={bsz} N={N} lower={lower}", flush=True)
    rng = np.random.default_rng(bsz * 7 + N)
    diag = rng.uniform(1.5, 2.5, (bsz, N))
    sdiag = rng.uniform(-1, 1, (bsz, N - 1))
    b = rng.normal(size=(bsz, N))
    got = rk.bidiagonal_solve_batched(_d(diag), _d(sdiag), _d(b), lower)
    torch.cuda.synchronize()
    _close(got, bd.bidiagonal_solve(diag, sdiag, b, lower))


@pytest.mark.parametrize("nd", [1, 2, 3, 4, 6, 8])
@pytest.mark.parametrize("lower", [True, False])
def test_banded_solve_and_mv(nd, lower):
    from boltzmann_amd import rawkernel as rk
    print(f"case: banded nd={nd} lower={lower}", flush=True)
    rng = np.random.default_rng(nd * 3 + lower)
    bsz, N = 150, 77
    # diagonally dominant, so the sweeps do not amplify rounding differences
    diags = rng.uniform(-1, 1, (bsz, nd, N)) / nd
    diags[:, 0, :] = rng.uniform(2.0, 3.0, (bsz, N))
    b = rng.normal(size=(bsz, N))
    x = rk.banded_tri_solve_batched(_d(diags), _d(b), lower)
    y = rk.banded_tri_mv_batched(_d(diags), _d(b), lower)
    torch.cuda.synchronize()
    _close(x, bd.banded_tri_solve(diags, b, lower))
    _close(y, bd.banded_tri_mv(diags, b, lower))


@pytest.mark.parametrize("kind", ["vr", "vt", "x"])
@pytest.mark.parametrize("nd", [2, 3])
def test_advection_sweeps(kind, nd):
    from boltzmann_amd import rawkernel as rk
    print(f"case: adv_{kind} nd={nd}", flush=True)
    rng = np.random.default_rng(11 + nd)
    Nr, Nvt, Nx, dt = 129, 32, 200, 0.05        # the 1D3V shapes of the glow case
    Ex = rng.normal(size=Nx) * 3
    cos_vt = np.cos(np.linspace(0.05, np.pi - 0.05, Nvt))
    vr = rng.uniform(0.2, 4.0, Nr)
    N = {"vr": Nr, "vt": Nvt, "x": Nx}[kind]

    def upwind(sign):
        # diagonal of the upwind sign (1 + a dt D_ii >= 1 on the systems that use it),
        # smaller neighbours of the other sign: a stable backward-Euler sweep
        D = -sign * rng.uniform(0, 0.5 / nd, (nd, N))
        D[0] = sign * rng.uniform(0.5, 1.0, N)
        return D
    # vr, x sweep left to right for a >= 0; vt for a <= 0
    DL, DR = (upwind(-1.0), upwind(1.0)) if kind == "vt" else (upwind(1.0), upwind(-1.0))
    if kind == "vr":
        b = rng.normal(size=(Nx * Nvt, Nr))
        ref = bd.adv_vr_BE(DL, DR, Ex, cos_vt, dt, b, Nr, Nvt, Nx)
        out = torch.empty_like(_d(b))
        rk.adv_vr_BE(_d(DL), _d(DR), _d(Ex), _d(cos_vt), dt, _d(b), out, nd, Nr, Nvt, Nx)
    elif kind == "vt":
        b = rng.normal(size=(Nx * Nr, Nvt))
        ref = bd.adv_vt_BE(DL, DR, Ex, vr, dt, b, Nr, Nvt, Nx)
        out = torch.empty_like(_d(b))
        rk.adv_vt_BE(_d(DL), _d(DR), _d(Ex), _d(vr), dt, _d(b), out, nd, Nr, Nvt, Nx)
    else:
        b = rng.normal(size=(Nr * Nvt, Nx))
        ref = bd.adv_x_BE(DL, DR, vr, cos_vt, dt, b, Nr, Nvt, Nx)
        out = torch.empty_like(_d(b))
        rk.adv_x_BE(_d(DL), _d(DR), _d(vr), _d(cos_vt), dt, _d(b), out, nd, Nr, Nvt, Nx)
    torch.cuda.synchronize()
    _close(out, ref)


@pytest.mark.parametrize("bsz,nrhs", [(4, 3), (32, 200), (64, 5)])
def test_bsr_upper_unit_solve(bsz, nrhs):
    from boltzmann_amd.rawkernel import BSRUpperUnitSolver
    print(f"case: bsr bsz={bsz} nrhs={nrhs}", flush=True)
    rng = np.random.default_rng(bsz + nrhs)
    nbr = 20
    indptr, indices, data = [0], [], []
    for i in range(nbr):
        for j in range(max(0, i - 1), min(nbr, i + 3)):
            indices.append(j)
            data.append(rng.normal(size=(bsz, bsz)) * (0.3 / bsz ** 0.5))
        indptr.append(len(indices))
    data = np.array(data)
    indptr = np.array(indptr, dtype=np.int32)
    indices = np.array(indices, dtype=np.int32)
    b = rng.normal(size=(nbr * bsz, nrhs))
    ref = bd.bsr_upper_unit_solve(data, indptr, indices, b)
    s = BSRUpperUnitSolver(_d(data), torch.from_numpy(indptr).cuda(), torch.from_numpy(indices).cuda())
    got = s.solve(_d(b))
    torch.cuda.synchronize()
    _close(got, ref)


def test_spsm_solver_mirror():
    """cusparse_spsolve.SpSMSolver's interface (scipy BSR matrix in, repeated solve(b)) over
    the BSR kernel: the unit-diagonal block upper triangle of a random block matrix against
    scipy.linalg.solve_triangular on the dense upper triangle with unit diagonal."""
    import scipy.linalg
    import scipy.sparse
    from boltzmann_amd.cusparse_spsolve import SpSMSolver, solver_type
    print("case: SpSMSolver mirror bsz=8 nbr=16", flush=True)
    rng = np.random.default_rng(9)
    bsz, nbr = 8, 16
    A = scipy.sparse.random(nbr * bsz, nbr * bsz, density=0.2, random_state=3, format="csr") * 0.2
    A = scipy.sparse.bsr_matrix(A.toarray() + np.eye(nbr * bsz), blocksize=(bsz, bsz))
    s = SpSMSolver(A, nrhs=4, lower=False, unit_diagonal=True, stype=solver_type.CUDA_CODE)
    # the kernel uses the blocks strictly above the diagonal; the diagonal blocks are identity
    dense = A.toarray()
    U = np.eye(nbr * bsz)
    for i in range(nbr):
        for j in range(i + 1, nbr):
            U[i * bsz:(i + 1) * bsz, j * bsz:(j + 1) * bsz] = dense[i * bsz:(i + 1) * bsz, j * bsz:(j + 1) * bsz]
    for nrhs in (1, 4):
        b = rng.normal(size=(nbr * bsz, nrhs)) if nrhs > 1 else rng.normal(size=nbr * bsz)
        ref = scipy.linalg.solve_triangular(U, b, lower=False, unit_diagonal=True)
        _close(s.solve(b), ref)
    with pytest.raises(NotImplementedError):
        SpSMSolver(A, lower=True)
