"""Oracle: batched bidiagonal / banded triangular solves and products, and the
backward-Euler upwind advection sweeps of the 1D3V solver (FP64, numpy).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates the raw kernels of BESolver/python/rawkernel.py, one system per row,
with the reference's operation order per element (vectorised over systems):
  * lower/upper_bidiagonal_batched       rawkernel.py:6-46
  * lower/upper_banded_tri_batched       rawkernel.py:48-109
  * lower/upper_banded_tri_batched_mv    rawkernel.py:111-166
  * adv_vr_BE_solve / adv_vt_BE_solve / adv_x_BE_solve   rawkernel.py:169-375
  * the BSR block upper-triangular unit-diagonal solve
    BESolver/cuda/spsm_bsr_upper_tri_solve.cu:2-58 (cusparse_spsolve.py:330-353)
Layouts: ``diag (bsz, N)``, ``subdiag (bsz, N-1)``, ``diags (bsz, ndiags, N)`` with
diagonal d of row i at ``[d, i]``; the advection operators ``D (ndiags, N)`` are
shared by all systems with entry ``D[d, i-d]`` (lower, left to right) or
``D[d, j+d]`` (upper, right to left) multiplying the neighbour d away.

Pinned by tests/test_oracle_banded.py: every solve against
scipy.linalg.solve_triangular on the dense matrix built from the diagonals, every
product against the dense product, the advection sweeps against the dense
(I + dt a D) solve, the BSR solve against a dense block solve.
"""
import numpy as np


def bidiagonal_solve(diag, sdiag, b, lower):
    """rawkernel.py:6-46."""
    bsz, N = b.shape
    out = np.zeros_like(b)
    if lower:
        out[:, 0] = b[:, 0] / diag[:, 0]
        for i in range(1, N):
            out[:, i] = (b[:, i] - sdiag[:, i - 1] * out[:, i - 1]) / diag[:, i]
    else:
        out[:, N - 1] = b[:, N - 1] / diag[:, N - 1]
        for i in range(1, N):
            j = N - 1 - i
            out[:, j] = (b[:, j] - sdiag[:, j] * out[:, j + 1]) / diag[:, j]
    return out


def banded_tri_solve(diags, b, lower):
    """rawkernel.py:48-109."""
    bsz, nd, N = diags.shape
    out = np.zeros_like(b)
    if lower:
        out[:, 0] = b[:, 0] / diags[:, 0, 0]
        for i in range(1, N):
            s = np.zeros(bsz)
            for d in range(1, nd):
                if i - d < 0:
                    break
                s = s + diags[:, d, i] * out[:, i - d]
            out[:, i] = (b[:, i] - s) / diags[:, 0, i]
    else:
        out[:, N - 1] = b[:, N - 1] / diags[:, 0, N - 1]
        for i in range(1, N):
            j = N - 1 - i
            s = np.zeros(bsz)
            for d in range(1, nd):
                if j + d >= N:
                    break
                s = s + diags[:, d, j] * out[:, j + d]
            out[:, j] = (b[:, j] - s) / diags[:, 0, j]
    return out


def banded_tri_mv(diags, b, lower):
    """rawkernel.py:111-166."""
    bsz, nd, N = diags.shape
    out = np.zeros_like(b)
    for i in range(N):
        s = np.zeros(bsz)
        for d in range(nd):
            if lower:
                if i - d < 0:
                    break
                s = s + diags[:, d, i] * b[:, i - d]
            else:
                if i + d >= N:
                    break
                s = s + diags[:, d, i] * b[:, i + d]
        out[:, i] = s
    return out


def _adv_sweep(DL, DR, av, dt, b, order):
    """One upwind backward-Euler sweep per system (rows of b), direction by sign.
    order = (forward grouping, backward grouping, forward mask); the reference groups
    the neighbour term as 'a' -> av*dt*s, 'b' -> s*av*dt or 'c' -> s*dt*av."""
    nd, N = DL.shape
    out = np.zeros_like(b)
    fwd = np.asarray(order[2], dtype=bool)       # systems swept left to right
    f = fwd
    if f.any():
        out[f, 0] = b[f, 0] / (1.0 + dt * av[f] * DL[0, 0])
        for i in range(1, N):
            s = np.zeros(int(f.sum()))
            for d in range(1, nd):
                if i - d < 0:
                    break
                s = s + DL[d, i - d] * out[f, i - d]
            avf = av[f]
            t = {"a": avf * dt * s, "b": s * avf * dt, "c": s * dt * avf}[order[0]]
            out[f, i] = (b[f, i] - t) / (1.0 + avf * dt * DL[0, i])
    r = ~fwd
    if r.any():
        avr = av[r]
        out[r, N - 1] = b[r, N - 1] / (1.0 + avr * dt * DR[0, N - 1])
        for i in range(1, N):
            j = N - 1 - i
            s = np.zeros(int(r.sum()))
            for d in range(1, nd):
                if j + d >= N:
                    break
                s = s + DR[d, j + d] * out[r, j + d]
            t = {"a": avr * dt * s, "b": s * avr * dt, "c": s * dt * avr}[order[1]]
            out[r, j] = (b[r, j] - t) / (1.0 + avr * dt * DR[0, j])
    return out


def adv_vr_BE(DL, DR, Ex, cos_vt, dt, b, Nr, Nvt, Nx):
    """rawkernel.py:169-231: b, out (Nx*Nvt, Nr); a = -E_x cos(theta)."""
    av = (-Ex[:, None] * cos_vt[None, :]).reshape(-1)
    return _adv_sweep(DL, DR, av, dt, b.reshape(Nx * Nvt, Nr), ("a", "c", av >= 0))


def adv_vt_BE(DL, DR, Ex, vr, dt, b, Nr, Nvt, Nx):
    """rawkernel.py:234-298: b, out (Nx*Nr, Nvt); a = E_x / v_r, left to right when -a >= 0."""
    av = (Ex[:, None] / vr[None, :]).reshape(-1)
    return _adv_sweep(DL, DR, av, dt, b.reshape(Nx * Nr, Nvt), ("b", "c", -av >= 0))


def adv_x_BE(DL, DR, vr, cos_vt, dt, b, Nr, Nvt, Nx):
    """rawkernel.py:301-375: b, out (Nr*Nvt, Nx); a = v_r cos(theta)."""
    av = (vr[:, None] * cos_vt[None, :]).reshape(-1)
    return _adv_sweep(DL, DR, av, dt, b.reshape(Nr * Nvt, Nx), ("b", "c", av >= 0))


def bsr_upper_unit_solve(data, indptr, indices, b):
    """spsm_bsr_upper_tri_solve.cu:2-58: x = U^-1 b for a block upper-triangular BSR U
    with unit diagonal blocks (diagonal and lower blocks are skipped).  b (n, nrhs)."""
    nbr = len(indptr) - 1
    bsz = data.shape[1]
    x = np.array(b, dtype=np.float64, copy=True)
    for i in range(nbr - 1, -1, -1):
        bi = x[i * bsz:(i + 1) * bsz, :].copy()
        for idx in range(indptr[i], indptr[i + 1]):
            j = indices[idx]
            if j <= i:
                continue
            bi = bi - data[idx] @ x[j * bsz:(j + 1) * bsz, :]
        x[i * bsz:(i + 1) * bsz, :] = bi
    return x
