"""The reference's repeated sparse triangular solver, cusparse_spsolve.SpSMSolver
(cusparse_spsolve.py:64-395), over this library's BSR kernel.

Same constructor and solve(b).  The analysis of the reference (cuSPARSE descriptors) is
replaced by a one-time upload and structure check of the matrix; every solve is one launch
of bz_bsr_upper_unit_solve (the CUDA_CODE branch's kernel, cusparse_spsolve.py:366-372).
That kernel is what the fully implicit 1D3V preconditioner uses: block upper triangular,
unit diagonal blocks, BSR.  Other triangles (lower, non-unit diagonal) are not part of this
library and raise NotImplementedError whatever `stype` asks for."""
import numpy as np
import torch

from . import _lib
from .rawkernel import BSRUpperUnitSolver


class solver_type:
    CUSPARSE_SPSM = 0
    CUSPARSE_BSRSM2 = 1
    CUPY_SPSM_SEQ = 2
    CUPY_SPSM_PAR = 3
    CUDA_CODE = 4


class SpSMSolver:
    def __init__(self, A_sp, nrhs: int = 1, lower: bool = True, dtype=np.float64, unit_diagonal: bool = True,
                 stype=solver_type.CUSPARSE_BSRSM2):
        if np.dtype(dtype) != np.float64:
            raise NotImplementedError("float64 only, as the reference")
        if lower or not unit_diagonal:
            raise NotImplementedError("block upper triangular with unit diagonal blocks only")
        fmt = A_sp.format
        if fmt == "csr":
            A_sp = A_sp.tobsr(blocksize=(1, 1))
        elif fmt != "bsr":
            raise ValueError(f"Unsupported sparse format: {fmt}")
        self.lower, self.udiag, self.stype, self.format = lower, unit_diagonal, stype, fmt
        self.n = A_sp.shape[0]
        self.nnz = int(A_sp.nnz)
        dev = _lib.require_device()
        data = torch.from_numpy(np.ascontiguousarray(A_sp.data, dtype=np.float64)).to(dev)
        indptr = torch.from_numpy(np.asarray(A_sp.indptr, dtype=np.int32)).to(dev)
        indices = torch.from_numpy(np.asarray(A_sp.indices, dtype=np.int32)).to(dev)
        self._solver = BSRUpperUnitSolver(data, indptr, indices)
        self.bsz = self._solver.bsz

    def solve(self, b):
        """x with U x = b; b (n,) or (n, nrhs) (device tensor or host array; a host array
        is copied to the device).  Returns a device tensor of b's shape."""
        if not isinstance(b, torch.Tensor):
            b = torch.from_numpy(np.ascontiguousarray(b, dtype=np.float64)).to(self._solver.data.device)
        return self._solver.solve(b)
