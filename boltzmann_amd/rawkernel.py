"""Batched triangular sweeps of the 1D3V solver, with the reference's Python interface.

Mirrors BESolver/python/rawkernel.py:389-503 (bidiagonal_solve_batched,
banded_tri_solve_batched, banded_tri_mv_batched, adv_vr_BE, adv_vt_BE, adv_x_BE)
over the HIP kernels of bz_banded.hip.  Arrays are float64 torch tensors on the
HIP device (the reference takes CuPy arrays); shapes and leading-dimension
flattening follow the reference wrappers, and the shape checks raise
AssertionError as the reference's asserts do.
"""
import torch

from . import _lib


def _c(t):
    assert isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64, "float64 device tensor expected"
    return t.contiguous()


def bidiagonal_solve_batched(diag, sdiag, b, lower):
    """rawkernel.py:389-413."""
    _diag = _c(diag).reshape((-1, diag.shape[-1]))
    _sdiag = _c(sdiag).reshape((-1, sdiag.shape[-1]))
    _b = _c(b).reshape((-1, b.shape[-1]))
    assert _sdiag.shape[0] == _diag.shape[0] and _sdiag.shape[0] == _b.shape[0]
    assert _sdiag.shape[1] == _diag.shape[1] - 1 and _sdiag.shape[1] == _b.shape[1] - 1
    bsz, N = _diag.shape
    out = torch.zeros_like(_b)
    _lib.call("bz_bidiag_solve_batched", _diag, _sdiag, _b, out, bsz, N, 1 if lower else 0)
    return out.reshape(diag.shape)


def banded_tri_solve_batched(diags, b, lower):
    """rawkernel.py:415-439."""
    _diags = _c(diags).reshape((-1, diags.shape[-2], diags.shape[-1]))
    _b = _c(b).reshape((-1, b.shape[-1]))
    assert _diags.shape[0] == _b.shape[0]
    bsz, nd, N = _diags.shape
    out = torch.zeros_like(_b)
    _lib.call("bz_banded_tri_solve_batched", _diags, _b, out, bsz, nd, N, 1 if lower else 0)
    return out.reshape(b.shape)


def banded_tri_mv_batched(diags, b, lower):
    """rawkernel.py:441-465."""
    _diags = _c(diags).reshape((-1, diags.shape[-2], diags.shape[-1]))
    _b = _c(b).reshape((-1, b.shape[-1]))
    assert _diags.shape[0] == _b.shape[0]
    bsz, nd, N = _diags.shape
    out = torch.zeros_like(_b)
    _lib.call("bz_banded_tri_mv_batched", _diags, _b, out, bsz, nd, N, 1 if lower else 0)
    return out.reshape(b.shape)


def adv_vr_BE(D_LtoR, D_RtoL, Ex, cos_vt, dt, b, out, ndiags, Nr, Nvt, Nx):
    """rawkernel.py:468-478: b, out (Nx*Nvt, Nr), Nr fastest."""
    assert out.is_contiguous()
    _lib.call("bz_adv_vr_be", _c(D_LtoR), _c(D_RtoL), _c(Ex), _c(cos_vt), float(dt), _c(b), out, int(ndiags),
              int(Nr), int(Nvt), int(Nx))


def adv_vt_BE(D_LtoR, D_RtoL, Ex, vr, dt, b, out, ndiags, Nr, Nvt, Nx):
    """rawkernel.py:480-490: b, out (Nx*Nr, Nvt), Nvt fastest."""
    assert out.is_contiguous()
    _lib.call("bz_adv_vt_be", _c(D_LtoR), _c(D_RtoL), _c(Ex), _c(vr), float(dt), _c(b), out, int(ndiags),
              int(Nr), int(Nvt), int(Nx))


def adv_x_BE(D_LtoR, D_RtoL, vr, cos_vt, dt, b, out, ndiags, Nr, Nvt, Nx):
    """rawkernel.py:492-503: b, out (Nr*Nvt, Nx), Nx fastest."""
    assert out.is_contiguous()
    _lib.call("bz_adv_x_be", _c(D_LtoR), _c(D_RtoL), _c(vr), _c(cos_vt), float(dt), _c(b), out, int(ndiags),
              int(Nr), int(Nvt), int(Nx))


class BSRUpperUnitSolver:
    """x = U^-1 b for a block upper-triangular BSR U with unit diagonal blocks: the
    CUDA_CODE branch of SpSMSolver (cusparse_spsolve.py:366-372) over
    spsm_bsr_upper_tri_solve.cu.  data (nnzb, bsz, bsz), indptr, indices on the device."""

    def __init__(self, data, indptr, indices):
        self.data = _c(data)
        self.indptr = indptr.to(torch.int32).contiguous()
        self.indices = indices.to(torch.int32).contiguous()
        self.bsz = int(data.shape[1])
        self.nbr = int(indptr.shape[0]) - 1
        self.n = self.nbr * self.bsz
        # the kernel trusts the structure: check it once on the host
        ip = self.indptr.cpu()
        assert data.dim() == 3 and data.shape[1] == data.shape[2] and self.nbr >= 1
        assert int(ip[0]) == 0 and bool((ip[1:] >= ip[:-1]).all())
        assert int(ip[-1]) == self.indices.numel() == data.shape[0]
        if self.indices.numel():
            assert int(self.indices.min()) >= 0 and int(self.indices.max()) < self.nbr

    def solve(self, b):
        """b (n, nrhs) or (n,); returns x with the same shape (C order)."""
        vec = b.dim() == 1
        bb = _c(b).reshape(self.n, -1)
        assert bb.shape[0] == self.n
        nrhs = bb.shape[1]
        bf = bb.t().contiguous()                  # column-major (n, nrhs) = row-major (nrhs, n)
        xf = torch.empty_like(bf)
        _lib.call("bz_bsr_upper_unit_solve", self.data, self.indptr, self.indices, self.nbr, nrhs, self.bsz, bf, xf)
        x = xf.t().contiguous()
        return x.reshape(-1) if vec else x
