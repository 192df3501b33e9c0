// The fully implicit 1D3V step on the finite-volume / ordinate velocity grid (gfx950, FP64):
// the step operator (rhs_mvec) and the XVTVRC operator-split preconditioner of
// bte_1d3v.step (BESolver/python/bte_1d3v_solver.py:1745-1931).
//
// State v (Nr*Nvt, Nx), x fastest, kept in that layout throughout: every kernel maps
// consecutive lanes to consecutive x points, so the radial and angular neighbours of an
// element are whole coalesced rows Nvt*Nx or Nx doubles away, and no transposes are needed
// (the reference swaps axes between its sweeps).
//
// The collision operator has the structure of the strong form: the gain of point (r, t) is
// (G s)[r], with s[r'] = sum_t 2 pi w_t v[r', t] the isotropic moment and G (Nr x Nr) at most
// two entries per process per row (linear interpolation at the pre-collision speed); the loss
// is diagonal.  So C v costs a moment pass and a short sparse row per element, and the
// collision solve (I - c C)^-1 b = b / A + H s_b (A = 1 + c loss, H a dense Nr x Nr matrix
// built once) costs a moment pass and one Nr x Nr by Nr x Nx product, instead of the
// reference's dense (Nr Nvt)^2 PmatC.
#include "bz_common.h"
#include "bz.h"

namespace {

constexpr int MAXB = 8;       // widest stencil (pvr, pvt, px <= 7)

// s[r, x] = sum_t wq[t] v[(r Nvt + t) Nx + x]
__global__ __launch_bounds__(256) void k_fvm_moment(const double* __restrict__ v, const double* __restrict__ wq,
                                                    int Nr, int Nvt, int Nx, double* __restrict__ s) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= Nr * Nx) return;
  const int r = e / Nx, ix = e - r * Nx;
  const double* p = v + (size_t)r * Nvt * Nx + ix;
  double acc = 0.0;
  for (int t = 0; t < Nvt; ++t) acc += wq[t] * p[(size_t)t * Nx];
  s[e] = acc;
}

// One element (r, t, x) of the step operator per thread.
__global__ __launch_bounds__(256) void k_fvm_matvec(BzFvm1d3v op, const double* __restrict__ E, double dt,
                                                    const double* __restrict__ x, const double* __restrict__ s,
                                                    double* __restrict__ y) {
  const int Nr = op.Nr, Nvt = op.Nvt, Nx = op.Nx;
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)Nr * Nvt * Nx) return;
  const int ix = (int)(e % Nx);
  const int rt = (int)(e / Nx);
  const int t = rt % Nvt, r = rt / Nvt;
  const int mp = Nvt / 2;
  const double cs = op.cos_vt[t];
  const bool low = t >= mp;                    // theta <= pi/2: cos > 0, left-to-right in x
  // x shift: sum_k (delta_k0 + dt_x adx_r cos_t D[k, ix]) x[ix -/+ k]
  const double cx = op.dt_x * op.adx[r] * cs;
  double yx = 0.0;
  for (int k = 0; k < op.nx; ++k) {
    const int j = low ? ix - k : ix + k;
    if (j < 0 || j >= Nx) break;
    const double d = cx * (low ? op.dxL : op.dxR)[k * Nx + ix] + (k == 0 ? 1.0 : 0.0);
    yx += d * x[e + (j - ix)];
  }
  const bool inflow = low ? (ix == 0) : (ix == Nx - 1);
  double out = yx;
  if (!inflow) {
    // collisions: (G s)[r] - loss[r] x
    double g = 0.0;
    for (int q = op.g_ptr[r]; q < op.g_ptr[r + 1]; ++q) g += op.g_val[q] * s[(size_t)op.g_col[q] * Nx + ix];
    const double xc = x[e];
    const double cv = g - op.loss[r] * xc;
    // advection: cos d/dvr - (sin / vr) d/dtheta, upwinded by the sign of E (Ep / En)
    const double Ex = E[ix];
    const bool ep = Ex >= 0;
    const bool vr_up = (ep == low);             // Dvr_RtoL: neighbours r + d
    const double* Dr = vr_up ? op.dvrR : op.dvrL;
    const long rs = (long)Nvt * Nx;
    double ar = 0.0;
    for (int d = 0; d < op.nvr; ++d) {
      const int j = vr_up ? r + d : r - d;
      if (j < 0 || j >= Nr) break;
      ar += Dr[d * Nr + r] * x[e + (vr_up ? d : -d) * rs];
    }
    const double* Dt = ep ? op.dvtR : op.dvtL;   // Ep: Dvt_RtoL (neighbours t + d)
    double at = 0.0;
    for (int d = 0; d < op.nvt; ++d) {
      const int j = ep ? t + d : t - d;
      if (j < 0 || j >= Nvt) break;
      at += Dt[d * Nvt + t] * x[e + (ep ? d : -d) * (long)Nx];
    }
    const double ae = cs * ar - (op.sin_vt[t] / op.vr[r]) * at;
    out = yx - dt * (op.tau_n0np0 * cv + op.tau * (Ex * (op.eq * ae)));
  }
  y[e] = out;
}

// h[r, x] = sum_{r' >= h_lo[r]} H[r, r'] s[r', x]; RB rows per thread
constexpr int RB = 4;
__global__ __launch_bounds__(256) void k_fvm_coll_corr(const double* __restrict__ H, const int* __restrict__ h_lo,
                                                       const double* __restrict__ s, int Nr, int Nx,
                                                       double* __restrict__ h) {
  const int nrb = (Nr + RB - 1) / RB;
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= nrb * Nx) return;
  const int rb = e / Nx, ix = e - rb * Nx;
  const int r0 = rb * RB;
  int lo = Nr;
  for (int j = 0; j < RB; ++j)
    if (r0 + j < Nr) lo = min(lo, h_lo[r0 + j]);
  double acc[RB];
#pragma unroll
  for (int j = 0; j < RB; ++j) acc[j] = 0.0;
  for (int c = lo; c < Nr; ++c) {
    const double sv = s[(size_t)c * Nx + ix];
#pragma unroll
    for (int j = 0; j < RB; ++j)
      if (r0 + j < Nr) acc[j] += H[(size_t)(r0 + j) * Nr + c] * sv;
  }
#pragma unroll
  for (int j = 0; j < RB; ++j)
    if (r0 + j < Nr) h[(size_t)(r0 + j) * Nx + ix] = acc[j];
}

// Backward-Euler upwind sweep along r (VR) or t (!VR) of one system per thread, in the state
// layout: VR systems (t, x), elements Nvt*Nx apart; angular systems (r, x), elements Nx apart.
// Same recurrence and operation order as rawkernel.adv_vr_BE_solve / adv_vt_BE_solve
// (rawkernel.py:169-298): out_c = (b_c - av dt sum_d D[d, c-/+d] out_(c-/+d)) /
// (1 + av dt D[0, c]), direction by the sign of the speed.  The radial sweep reads the
// collision solve's output b / A + h on the fly.
template <bool VR, int ND>
__global__ __launch_bounds__(256) void k_fvm_sweep(BzFvm1d3v op, const double* __restrict__ E, double dt,
                                                   const double* b, double* out,   // may alias
                                                   const double* __restrict__ h) {
  const int Nr = op.Nr, Nvt = op.Nvt, Nx = op.Nx;
  const int s = blockIdx.x * 256 + threadIdx.x;
  const int nsys = VR ? Nvt * Nx : Nr * Nx;
  if (s >= nsys) return;
  const int ix = s % Nx, o = s / Nx;            // o = t (VR) or r
  const double aE = op.eaE * E[ix];
  double av;
  bool fwd;
  if (VR) { av = -aE * op.cos_vt[o]; fwd = av >= 0; }
  else { av = aE / op.vr[o]; fwd = -av >= 0; }
  const int N = VR ? Nr : Nvt;
  const long base = VR ? (long)o * Nx + ix : (long)o * Nvt * Nx + ix;
  const long stride = VR ? (long)Nvt * Nx : (long)Nx;
  const double* D = VR ? (fwd ? op.vr_l2r : op.vr_r2l) : (fwd ? op.vt_l2r : op.vt_r2l);
  double y[ND > 1 ? ND - 1 : 1];
#pragma unroll
  for (int d = 0; d < (ND > 1 ? ND - 1 : 1); ++d) y[d] = 0.0;
  for (int p = 0; p < N; ++p) {
    const int c = fwd ? p : N - 1 - p;
    const long idx = base + c * stride;
    const double bi = VR ? b[idx] * op.ainv[c] + h[(size_t)c * Nx + ix] : b[idx];
    double sv = 0.0;
#pragma unroll
    for (int d = 1; d < ND; ++d)
      if (d <= p) sv += D[d * N + (fwd ? c - d : c + d)] * y[d - 1];
    const double term = !fwd ? sv * dt * av : (VR ? av * dt * sv : sv * av * dt);
    const double val = (bi - term) / (1.0 + av * dt * D[c]);
#pragma unroll
    for (int d = (ND > 1 ? ND - 2 : 0); d >= 1; --d) y[d] = y[d - 1];
    y[0] = val;
    out[idx] = val;
  }
}

template <bool VR>
int launch_fvm_sweep(const BzFvm1d3v& op, const double* E, double dt, const double* b, double* out, const double* h,
                     hipStream_t st) {
  const int nsys = VR ? op.Nvt * op.Nx : op.Nr * op.Nx;
  const int nd = VR ? op.nvr : op.nvt;
  const dim3 grid((nsys + 255) / 256), block(256);
  switch (nd) {
    case 1: hipLaunchKernelGGL((k_fvm_sweep<VR, 1>), grid, block, 0, st, op, E, dt, b, out, h); break;
    case 2: hipLaunchKernelGGL((k_fvm_sweep<VR, 2>), grid, block, 0, st, op, E, dt, b, out, h); break;
    case 3: hipLaunchKernelGGL((k_fvm_sweep<VR, 3>), grid, block, 0, st, op, E, dt, b, out, h); break;
    case 4: hipLaunchKernelGGL((k_fvm_sweep<VR, 4>), grid, block, 0, st, op, E, dt, b, out, h); break;
    case 5: hipLaunchKernelGGL((k_fvm_sweep<VR, 5>), grid, block, 0, st, op, E, dt, b, out, h); break;
    case 6: hipLaunchKernelGGL((k_fvm_sweep<VR, 6>), grid, block, 0, st, op, E, dt, b, out, h); break;
    case 7: hipLaunchKernelGGL((k_fvm_sweep<VR, 7>), grid, block, 0, st, op, E, dt, b, out, h); break;
    case 8: hipLaunchKernelGGL((k_fvm_sweep<VR, 8>), grid, block, 0, st, op, E, dt, b, out, h); break;
    default: return BZ_EINVAL;
  }
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int check_op(const BzFvm1d3v* op) {
  if (!op) return BZ_EINVAL;
  if (op->Nr < 2 || op->Nvt < 2 || (op->Nvt & 1) || op->Nx < 2) return BZ_EINVAL;
  if (op->nvr < 1 || op->nvr > MAXB || op->nvt < 1 || op->nvt > MAXB || op->nx < 1 || op->nx > MAXB) return BZ_EINVAL;
  if ((long)op->Nr * op->Nvt * op->Nx > (1L << 31) - 256) return BZ_EINVAL;
  if (!op->work) return BZ_EINVAL;
  return BZ_OK;
}

}  // namespace

extern "C" {

int bz_fvm_rhs_mvec(const BzFvm1d3v* op, const double* E, double dt, const double* x, double* y,
                    hipStream_t stream) {
  int rc = check_op(op);
  if (rc != BZ_OK) return rc;
  if (!E || !x || !y || x == y) return BZ_EINVAL;
  if (!op->cos_vt || !op->sin_vt || !op->vr || !op->wq || !op->loss || !op->g_ptr || !op->g_col || !op->g_val ||
      !op->dvrL || !op->dvrR || !op->dvtL || !op->dvtR || !op->dxL || !op->dxR || !op->adx)
    return BZ_EINVAL;
  const int nrx = op->Nr * op->Nx;
  double* s = op->work;
  hipLaunchKernelGGL(k_fvm_moment, dim3((nrx + 255) / 256), dim3(256), 0, stream, x, op->wq, op->Nr, op->Nvt, op->Nx,
                     s);
  BZ_LAUNCH_CHECK();
  const long n = (long)nrx * op->Nvt;
  hipLaunchKernelGGL(k_fvm_matvec, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, *op, E, dt, x, s, y);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_fvm_precond_xvtvrc(const BzFvm1d3v* op, const double* E, double alpha_dt, const double* b, double* y,
                          hipStream_t stream) {
  int rc = check_op(op);
  if (rc != BZ_OK) return rc;
  if (!E || !b || !y || b == y) return BZ_EINVAL;
  if (!op->cos_vt || !op->vr || !op->wq || !op->vr_l2r || !op->vr_r2l || !op->vt_l2r || !op->vt_r2l ||
      !op->x_l2r || !op->x_r2l || !op->axvr || !op->ainv || !op->H || !op->h_lo)
    return BZ_EINVAL;
  const int nrx = op->Nr * op->Nx;
  double* s = op->work;
  double* h = op->work + nrx;
  // collisions: s_b, then h = H s_b (the sweep adds b / A + h)
  hipLaunchKernelGGL(k_fvm_moment, dim3((nrx + 255) / 256), dim3(256), 0, stream, b, op->wq, op->Nr, op->Nvt, op->Nx,
                     s);
  BZ_LAUNCH_CHECK();
  const int nrb = (op->Nr + RB - 1) / RB;
  hipLaunchKernelGGL(k_fvm_coll_corr, dim3((nrb * op->Nx + 255) / 256), dim3(256), 0, stream, op->H, op->h_lo, s,
                     op->Nr, op->Nx, h);
  BZ_LAUNCH_CHECK();
  rc = launch_fvm_sweep<true>(*op, E, alpha_dt, b, y, h, stream);
  if (rc != BZ_OK) return rc;
  rc = launch_fvm_sweep<false>(*op, E, alpha_dt, y, y, nullptr, stream);
  if (rc != BZ_OK) return rc;
  return bz_internal_adv_x_be_inflow0(op->x_l2r, op->x_r2l, op->axvr, op->cos_vt, alpha_dt, y, y, op->nx, op->Nr,
                                      op->Nvt, op->Nx, stream);
}

}  // extern "C"
