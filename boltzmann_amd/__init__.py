"""MI355X-native velocity-space hot path of the B-spline x spherical-harmonic
electron Boltzmann solver (drop-in for the reference's batched 0D solver,
operator assembly and 1D3V velocity-space kernels).

Compute runs in hand-written HIP kernels (libbz.so, C ABI in include/bz.h);
PyTorch provides device memory and streams only.
"""
__all__ = ["build"]
