"""Functions of the reference's utils module (BESolver/python/utils.py) that the v-space path
calls, with the reference's signatures, over this library's device implementations."""
from .krylov import gmres_pcr as _gmres_pcr


def gmres_pcr(xp, A, b, x0=None, atol=1e-20, rtol=1e-10, maxiter=None, restart=None, M=None, callback=None):
    """utils.gmres_pcr (utils.py:1172-1268): restarted right-preconditioned GMRES.  `xp` (the
    reference's array module) is accepted and ignored: b, x0 and the operands are torch
    device tensors and every step runs on the HIP device (boltzmann_amd.krylov)."""
    return _gmres_pcr(A, b, x0=x0, atol=atol, rtol=rtol, maxiter=maxiter, restart=restart, M=M, callback=callback)
