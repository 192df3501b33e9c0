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
  // eight ordinates' loads in flight at a time (the sum keeps the ordinate order)
  double acc = 0.0;
  int t = 0;
  for (; t + 8 <= Nvt; t += 8) {
    double q[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) q[j] = p[(size_t)(t + j) * Nx];
#pragma unroll
    for (int j = 0; j < 8; ++j) acc += wq[t + j] * q[j];
  }
  for (; t < Nvt; ++t) acc += wq[t] * p[(size_t)t * Nx];
  s[e] = acc;
}

// One element (r, t, x) of the step operator per thread.
__global__ __launch_bounds__(256) void k_fvm_matvec(BzFvm1d3v op, const double* __restrict__ E, double dt,
                                                    const double* __restrict__ x, const double* __restrict__ s,
                                                    double* __restrict__ y) {
  const int Nr = op.Nr, Nvt = op.Nvt, Nx = op.Nx;
  // 32-bit index arithmetic (the size is checked below 2^31 by the host)
  const unsigned eu = blockIdx.x * 256u + threadIdx.x;
  if (eu >= (unsigned)(Nr * Nvt * Nx)) return;
  const long e = eu;
  const unsigned rt = eu / (unsigned)Nx;
  const int ix = (int)(eu - rt * (unsigned)Nx);
  const int r = (int)(rt / (unsigned)Nvt), t = (int)(rt - (unsigned)r * (unsigned)Nvt);
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

// The step operator for fixed stencil widths (the common ones): every neighbour load is issued
// unconditionally from a clamped address and weighted by zero where the stencil has no entry,
// so all loads of an element go out together instead of one loop iteration at a time; the
// gain row is read the same way, GMAX entries at once.  Same sums, in the same order, as
// k_fvm_matvec (the zero terms add exactly nothing).
template <int NX, int NVR, int NVT, int GMAX>
__global__ __launch_bounds__(256) void k_fvm_matvec_w(BzFvm1d3v op, const double* __restrict__ E, double dt,
                                                      const double* __restrict__ x, const double* __restrict__ s,
                                                      double* __restrict__ y) {
  const int Nr = op.Nr, Nvt = op.Nvt, Nx = op.Nx;
  // Workgroups of one state row (r, t) each, threads over x: the row indices are workgroup-
  // uniform (no per-thread division).  Workgroups are dealt to the 8 XCDs in turn (workgroup
  // b to XCD b % 8); the grid is a multiple of 8 and XCD k takes the k-th eighth of the rows,
  // so the radial neighbours (Nvt rows away) of a row are read through the same L2.
  const int bx = (Nx + blockDim.x - 1) / blockDim.x;            // workgroups per row
  const unsigned per = gridDim.x >> 3;
  const unsigned lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const int row = (int)(lb / (unsigned)bx);
  const int ix = (int)(lb - (unsigned)row * bx) * blockDim.x + threadIdx.x;
  if (row >= Nr * Nvt || ix >= Nx) return;
  const int r = row / Nvt, t = row - r * Nvt;
  const long e = (long)row * Nx + ix;
  const int mp = Nvt / 2;
  const double cs = op.cos_vt[t];
  const bool low = t >= mp;
  const double cx = op.dt_x * op.adx[r] * cs;
  const double* DX = low ? op.dxL : op.dxR;
  const double x0 = x[e];
  double xv[NX], dx[NX];
#pragma unroll
  for (int k = 0; k < NX; ++k) {
    const int j = min(max(low ? ix - k : ix + k, 0), Nx - 1);
    xv[k] = k == 0 ? x0 : x[e + (j - ix)];
    dx[k] = DX[k * Nx + ix];
  }
  double yx = 0.0;
#pragma unroll
  for (int k = 0; k < NX; ++k) {
    const int j = low ? ix - k : ix + k;
    const double d = cx * dx[k] + (k == 0 ? 1.0 : 0.0);
    yx += (j >= 0 && j < Nx) ? d * xv[k] : 0.0;
  }
  const bool inflow = low ? (ix == 0) : (ix == Nx - 1);
  double out = yx;
  if (!inflow) {
    const int q0 = op.g_ptr[r], q1 = op.g_ptr[r + 1];
    const int qmax = max(op.g_ptr[Nr] - 1, 0);
    double gv[GMAX], gs[GMAX];
#pragma unroll
    for (int k = 0; k < GMAX; ++k) {
      const int q = min(q0 + k, qmax);
      gv[k] = op.g_val[q];
      gs[k] = s[(size_t)op.g_col[q] * Nx + ix];
    }
    double g = 0.0;
#pragma unroll
    for (int k = 0; k < GMAX; ++k) g += (q0 + k < q1) ? gv[k] * gs[k] : 0.0;
    for (int q = q0 + GMAX; q < q1; ++q) g += op.g_val[q] * s[(size_t)op.g_col[q] * Nx + ix];
    const double xc = x0;
    const double cv = g - op.loss[r] * xc;
    const double Ex = E[ix];
    const bool ep = Ex >= 0;
    const bool vr_up = (ep == low);
    const double* Dr = vr_up ? op.dvrR : op.dvrL;
    const long rs = (long)Nvt * Nx;
    double xr[NVR], dr[NVR];
#pragma unroll
    for (int d = 0; d < NVR; ++d) {
      const int j = min(max(vr_up ? r + d : r - d, 0), Nr - 1);
      xr[d] = d == 0 ? x0 : x[e + (long)(j - r) * rs];
      dr[d] = Dr[d * Nr + r];
    }
    double ar = 0.0;
#pragma unroll
    for (int d = 0; d < NVR; ++d) {
      const int j = vr_up ? r + d : r - d;
      ar += (j >= 0 && j < Nr) ? dr[d] * xr[d] : 0.0;
    }
    const double* Dt = ep ? op.dvtR : op.dvtL;
    double xt[NVT], dtt[NVT];
#pragma unroll
    for (int d = 0; d < NVT; ++d) {
      const int j = min(max(ep ? t + d : t - d, 0), Nvt - 1);
      xt[d] = d == 0 ? x0 : x[e + (long)(j - t) * Nx];
      dtt[d] = Dt[d * Nvt + t];
    }
    double at = 0.0;
#pragma unroll
    for (int d = 0; d < NVT; ++d) {
      const int j = ep ? t + d : t - d;
      at += (j >= 0 && j < Nvt) ? dtt[d] * xt[d] : 0.0;
    }
    const double ae = cs * ar - (op.sin_vt[t] / op.vr[r]) * at;
    out = yx - dt * (op.tau_n0np0 * cv + op.tau * (Ex * (op.eq * ae)));
  }
  y[e] = out;
}

// h = H s on the matrix cores: one workgroup of CC_WAVES wavefronts per 16 x 16 tile of h
// (rows r, points x).  The columns r' run from the first nonzero column of the tile's rows (H
// is upper triangular: collisions only move electrons down in energy) to Nr in steps of 4,
// split in CC_WAVES ranges, one per wavefront, each accumulated by v_mfma_f64_16x16x4f64; the
// partial tiles are summed through LDS.  Operand layout of the instruction (measured on
// gfx950): A (16 x 4) lane = i + 16 k, B (4 x 16) lane = j + 16 k; element m of the
// accumulator of lane l is D[4 m + l / 16][l % 16].  Rows and points beyond the ends are
// clamped on load and not stored.
typedef double bz_d4 __attribute__((ext_vector_type(4)));

constexpr int CC_WAVES = 8;

__global__ __launch_bounds__(64 * CC_WAVES) void k_fvm_coll_corr(const double* __restrict__ H,
                                                                 const int* __restrict__ h_lo,
                                                                 const double* __restrict__ s, int Nr, int Nx,
                                                                 double* __restrict__ h) {
  __shared__ double red[CC_WAVES - 1][256];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i16 = lane & 15, k4 = lane >> 4;
  const int r0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
  const int row = min(r0 + i16, Nr - 1), col = min(c0 + i16, Nx - 1);
  int kmin = h_lo[row];
#pragma unroll
  for (int o = 8; o >= 1; o >>= 1) kmin = min(kmin, __shfl_xor(kmin, o, 16));
  kmin = __shfl(kmin, 0, 64);
  const int k0 = kmin & ~3;
  const int nsteps = (Nr - k0 + 3) >> 2;
  const int per = (nsteps + CC_WAVES - 1) / CC_WAVES;
  const int s_lo = wave * per, s_hi = min(nsteps, s_lo + per);
  const double* Hr = H + (size_t)row * Nr;
  bz_d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int st = s_lo; st < s_hi; ++st) {
    const int kk = k0 + 4 * st + k4;
    const int kc = min(kk, Nr - 1);
    const double a = kk < Nr ? Hr[kc] : 0.0;
    const double bv = s[(size_t)kc * Nx + col];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc, 0, 0, 0);
  }
  if (wave > 0) {
#pragma unroll
    for (int m = 0; m < 4; ++m) red[wave - 1][m * 64 + lane] = acc[m];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      double v = acc[m];
#pragma unroll
      for (int w = 0; w < CC_WAVES - 1; ++w) v += red[w][m * 64 + lane];
      const int rr = r0 + 4 * m + k4, cc = c0 + i16;
      if (rr < Nr && cc < Nx) h[(size_t)rr * Nx + cc] = v;
    }
  }
}

// Backward-Euler upwind sweeps of the preconditioner, one system per thread, in the state
// layout: radial systems (t, x) with elements Nvt*Nx apart, angular systems (r, x) with
// elements Nx apart, x systems (r, t) contiguous.  The recurrence of
// rawkernel.adv_vr_BE_solve / adv_vt_BE_solve / adv_x_BE_solve (rawkernel.py:169-375):
// out_c = (b_c - av dt sum_d D[d, c-/+d] out_(c-/+d)) / (1 + av dt D[0, c]), direction by the
// sign of the speed; the division is a product with the reciprocal, formed ahead of the
// recurrence (the results differ from the division's in the last bit).  The radial sweep reads the collision solve's output b / A + h on the
// fly; the x sweep takes the inflow value of every system as zero.  The operands of the next
// PF positions are loaded ahead of the recurrence (they do not depend on it), and the
// shared band tables sit in LDS when they fit, so each step waits on arithmetic only.
enum { SW_VR = 0, SW_VT = 1, SW_X = 2 };
constexpr int SWT = 64;
constexpr int SW_LDS_DOUBLES = 8192;

template <int MODE, int ND, bool LDS>
__global__ __launch_bounds__(SWT) void k_fvm_sweep(BzFvm1d3v op, const double* __restrict__ E, double dt,
                                                   const double* b, double* out,   // may alias
                                                   const double* __restrict__ h) {
  extern __shared__ double sm[];
  constexpr int PF = (MODE == SW_VR && ND <= 4) ? 16 : 8;   // the radial sweep is the longest
  const int Nr = op.Nr, Nvt = op.Nvt, Nx = op.Nx;
  const int N = MODE == SW_VR ? Nr : (MODE == SW_VT ? Nvt : Nx);
  const double* DLt = MODE == SW_VR ? op.vr_l2r : (MODE == SW_VT ? op.vt_l2r : op.x_l2r);
  const double* DRt = MODE == SW_VR ? op.vr_r2l : (MODE == SW_VT ? op.vt_r2l : op.x_r2l);
  const double* ainv = op.ainv;
  if (LDS) {
    for (int i = threadIdx.x; i < ND * N; i += SWT) {
      sm[i] = DLt[i];
      sm[ND * N + i] = DRt[i];
    }
    if (MODE == SW_VR)
      for (int i = threadIdx.x; i < N; i += SWT) sm[2 * ND * N + i] = op.ainv[i];
    __syncthreads();
    DLt = sm;
    DRt = sm + ND * N;
    ainv = sm + 2 * ND * N;
  }
  const int s = blockIdx.x * SWT + threadIdx.x;
  const int nsys = MODE == SW_VR ? Nvt * Nx : (MODE == SW_VT ? Nr * Nx : Nr * Nvt);
  if (s >= nsys) return;
  double av;
  bool fwd;
  long base, stride;
  int ix = 0;
  if (MODE == SW_VR) {
    ix = s % Nx;
    const int t = s / Nx;
    const double aE = op.eaE * E[ix];
    av = -aE * op.cos_vt[t];
    fwd = av >= 0;
    base = (long)t * Nx + ix;
    stride = (long)Nvt * Nx;
  } else if (MODE == SW_VT) {
    ix = s % Nx;
    const int r = s / Nx;
    const double aE = op.eaE * E[ix];
    av = aE / op.vr[r];
    fwd = -av >= 0;
    base = (long)r * Nvt * Nx + ix;
    stride = Nx;
  } else {
    const int r = s / Nvt, t = s % Nvt;
    av = op.axvr[r] * op.cos_vt[t];
    fwd = av >= 0;
    base = (long)s * Nx;
    stride = 1;
  }
  const double* D = fwd ? DLt : DRt;
  const double dtav = dt * av;
  // operands of the next PF positions: b (and h), and the band entries of the position
  double pb[PF], ph[PF], pd[PF][ND];
  auto fetch = [&](int i, int pn) __attribute__((always_inline)) {
    const int cn = fwd ? pn : N - 1 - pn;
    pb[i] = b[base + cn * stride];
    if (MODE == SW_VR) ph[i] = h[(size_t)cn * Nx + ix];
#pragma unroll
    for (int d = 0; d < ND; ++d) pd[i][d] = D[d * N + (fwd ? max(cn - d, 0) : min(cn + d, N - 1))];
  };
#pragma unroll
  for (int i = 0; i < PF; ++i) fetch(i, min(i, N - 1));
  double y[ND > 1 ? ND - 1 : 1];
#pragma unroll
  for (int d = 0; d < (ND > 1 ? ND - 1 : 1); ++d) y[d] = 0.0;
  // One position of the sweep; the prefetch index is clamped (the last position is loaded
  // again near the end) and neighbours before the start of the sweep multiply y = 0 through
  // a clamped table index, so the body has no branches.
  auto step = [&](int i, int p) __attribute__((always_inline)) {
    const int c = fwd ? p : N - 1 - p;
    double bi = pb[i];
    if (MODE == SW_VR) bi = bi * ainv[c] + ph[i];
    if (MODE == SW_X) bi = (p == 0) ? 0.0 : bi;
    double dd[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) dd[d] = pd[i][d];
    fetch(i, min(p + PF, N - 1));
    // older neighbours first, so that only the last product waits on the previous value
    double sv = 0.0;
#pragma unroll
    for (int d = ND - 1; d >= 1; --d) sv += dd[d] * y[d - 1];
    const double rden = 1.0 / (1.0 + av * dt * dd[0]);   // off the recurrence
    const double val = (bi - sv * dtav) * rden;
#pragma unroll
    for (int d = (ND > 1 ? ND - 2 : 0); d >= 1; --d) y[d] = y[d - 1];
    y[0] = val;
    out[base + c * stride] = val;
  };
  const int nfull = N - N % PF;
  for (int p0 = 0; p0 < nfull; p0 += PF) {
#pragma unroll
    for (int i = 0; i < PF; ++i) step(i, p0 + i);
  }
#pragma unroll
  for (int i = 0; i < PF; ++i)
    if (nfull + i < N) step(i, nfull + i);
}

// The radial sweep as a segmented scan.  A radial system has Nr positions (hundreds) but
// there are only Nvt*Nx systems, so one thread per system leaves most of the chip idle and
// waits on a long chain.  Here each system is cut into NSEG segments, one thread each.  Along
// the sweep the state (the last K = ND-1 values) obeys an affine map per position,
// y_new = alpha - sum_d beta_d y_(d-1) (alpha = b rden, beta_d = av dt D_d rden); pass 1 composes
// the maps of its segment (K x K matrix A and offset c), the entering state of a segment is
// the composition of the earlier segments applied to the zero state the sweep starts from
// (through LDS), and pass 2 re-runs the segment with the sequential formula from that state.
// The maps are contractions (upwind, 1 + av dt D_0 >= 1 in the direction of the sweep), so
// the re-association changes the result at rounding level.  A workgroup holds SB systems
// (consecutive x) times NSEG segments, so loads of one segment are coalesced over x.
constexpr int SCAN_SB = 8, NSEG = 32;

// MAXL > 0: a segment has at most MAXL positions, and all of its loads (right-hand sides and
// stencil weights, from clamped in-bounds positions) are issued up front into registers and kept
// for pass 2, so each pass waits on memory once instead of once per position; the arithmetic,
// and its order, are those of the MAXL = 0 form, which loads one position at a time.
template <int ND, int MAXL, int SB>
__global__ __launch_bounds__(SB * NSEG) void k_fvm_vr_scan(BzFvm1d3v op, const double* __restrict__ E, double dt,
                                                        const double* __restrict__ b, double* __restrict__ out,
                                                        const double* __restrict__ h) {
  constexpr int K = ND > 1 ? ND - 1 : 1;
  constexpr int NL = MAXL > 0 ? MAXL : 1;
  __shared__ double maps[SB * NSEG][K * K + K];
  const int Nr = op.Nr, Nvt = op.Nvt, Nx = op.Nx;
  const int N = Nr;
  const int j = threadIdx.x % SB, seg = threadIdx.x / SB;
  const int s = blockIdx.x * SB + j;
  const bool active = s < Nvt * Nx;
  const int ss = active ? s : 0;
  const int ix = ss % Nx, t = ss / Nx;
  const double aE = op.eaE * E[ix];
  const double av = -aE * op.cos_vt[t];
  const bool fwd = av >= 0;
  const double dtav = dt * av;
  const long base = (long)t * Nx + ix, stride = (long)Nvt * Nx;
  const double* D = fwd ? op.vr_l2r : op.vr_r2l;
  const int L = (N + NSEG - 1) / NSEG;
  const int p_lo = min(seg * L, N), p_hi = min(p_lo + L, N);
  auto coeffs = [&](int p, double& bi, double dd[ND]) __attribute__((always_inline)) {
    const int c = fwd ? p : N - 1 - p;
    bi = b[base + c * stride] * op.ainv[c] + h[(size_t)c * Nx + ix];
#pragma unroll
    for (int d = 0; d < ND; ++d) dd[d] = D[d * N + (fwd ? max(c - d, 0) : min(c + d, N - 1))];
  };
  double bv[NL], dv[NL][ND];
  if constexpr (MAXL > 0) {
#pragma unroll
    for (int i = 0; i < MAXL; ++i) coeffs(min(p_lo + i, N - 1), bv[i], dv[i]);
  }
  // pass 1: the segment's map, state (y0 .. y_(K-1)), y0 the newest value
  double A[K][K], cv[K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    cv[r] = 0.0;
#pragma unroll
    for (int q = 0; q < K; ++q) A[r][q] = (r == q) ? 1.0 : 0.0;
  }
  auto compose = [&](double bi, const double dd[ND]) __attribute__((always_inline)) {
    const double rden = 1.0 / (1.0 + av * dt * dd[0]);
    double a0[K], c0 = bi * rden;
#pragma unroll
    for (int q = 0; q < K; ++q) a0[q] = 0.0;
#pragma unroll
    for (int d = ND - 1; d >= 1; --d) {
      const double beta = dtav * dd[d] * rden;
      c0 -= beta * cv[d - 1];
#pragma unroll
      for (int q = 0; q < K; ++q) a0[q] -= beta * A[d - 1][q];
    }
#pragma unroll
    for (int r = K - 1; r >= 1; --r) {
      cv[r] = cv[r - 1];
#pragma unroll
      for (int q = 0; q < K; ++q) A[r][q] = A[r - 1][q];
    }
    cv[0] = c0;
#pragma unroll
    for (int q = 0; q < K; ++q) A[0][q] = a0[q];
  };
  if (active) {
    if constexpr (MAXL > 0) {
#pragma unroll
      for (int i = 0; i < MAXL; ++i)
        if (p_lo + i < p_hi) compose(bv[i], dv[i]);
    } else {
      for (int p = p_lo; p < p_hi; ++p) {
        double bi, dd[ND];
        coeffs(p, bi, dd);
        compose(bi, dd);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < K; ++r) {
    maps[threadIdx.x][K * K + r] = cv[r];
#pragma unroll
    for (int q = 0; q < K; ++q) maps[threadIdx.x][r * K + q] = A[r][q];
  }
  __syncthreads();
  if (!active || p_lo >= p_hi) return;
  // entering state: the earlier segments' maps applied in order to the zero state
  double y[K];
#pragma unroll
  for (int r = 0; r < K; ++r) y[r] = 0.0;
  for (int sg = 0; sg < seg; ++sg) {
    const double* m = maps[sg * SB + j];
    double yn[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
      double acc = m[K * K + r];
#pragma unroll
      for (int q = 0; q < K; ++q) acc += m[r * K + q] * y[q];
      yn[r] = acc;
    }
#pragma unroll
    for (int r = 0; r < K; ++r) y[r] = yn[r];
  }
  // pass 2: the sequential recurrence over the segment
  auto advance = [&](int p, double bi, const double dd[ND]) __attribute__((always_inline)) {
    double sv = 0.0;
#pragma unroll
    for (int d = ND - 1; d >= 1; --d) sv += dd[d] * y[d - 1];
    const double rden = 1.0 / (1.0 + av * dt * dd[0]);
    const double val = (bi - sv * dtav) * rden;
#pragma unroll
    for (int d = K - 1; d >= 1; --d) y[d] = y[d - 1];
    y[0] = val;
    out[base + (fwd ? p : N - 1 - p) * stride] = val;
  };
  if constexpr (MAXL > 0) {
#pragma unroll
    for (int i = 0; i < MAXL; ++i)
      if (p_lo + i < p_hi) advance(p_lo + i, bv[i], dv[i]);
  } else {
    for (int p = p_lo; p < p_hi; ++p) {
      double bi, dd[ND];
      coeffs(p, bi, dd);
      advance(p, bi, dd);
    }
  }
}

template <int ND, int SB>
void launch_vr_scan(const BzFvm1d3v& op, const double* E, double dt, const double* b, double* out, const double* h,
                    hipStream_t st) {
  const int nsys = op.Nvt * op.Nx;
  const int L = (op.Nr + NSEG - 1) / NSEG;
  const dim3 grid((nsys + SB - 1) / SB), block(SB * NSEG);
  if (L <= 8) hipLaunchKernelGGL((k_fvm_vr_scan<ND, 8, SB>), grid, block, 0, st, op, E, dt, b, out, h);
  else if (L <= 24) hipLaunchKernelGGL((k_fvm_vr_scan<ND, 24, SB>), grid, block, 0, st, op, E, dt, b, out, h);
  else hipLaunchKernelGGL((k_fvm_vr_scan<ND, 0, SB>), grid, block, 0, st, op, E, dt, b, out, h);
}

int launch_fvm_vr_scan(const BzFvm1d3v& op, const double* E, double dt, const double* b, double* out,
                       const double* h, hipStream_t st) {
  switch (op.nvr) {
    case 2: launch_vr_scan<2, SCAN_SB>(op, E, dt, b, out, h, st); break;
    case 3: launch_vr_scan<3, SCAN_SB>(op, E, dt, b, out, h, st); break;
    case 4: launch_vr_scan<4, SCAN_SB>(op, E, dt, b, out, h, st); break;
    case 5: launch_vr_scan<5, SCAN_SB>(op, E, dt, b, out, h, st); break;
    default: return BZ_EINVAL;
  }
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

// The x sweep for two-point stencils (px = 1) as an affine scan, one wavefront per (r, t)
// system: along the sweep y_c = alpha_c + beta_c y_prev (alpha = b rden, beta = -av dt D_1 rden,
// the inflow position has alpha = beta = 0); each lane composes the maps of PER consecutive
// positions, a 6-step shuffle scan gives every lane the value entering its positions, and
// the lane re-runs them.  The maps are contractions (upwind), so the re-association changes
// the result at rounding level.
template <int PER>
__global__ __launch_bounds__(64) void k_fvm_x_scan(BzFvm1d3v op, double dt, const double* b, double* out) {
  const int s = blockIdx.x, lane = threadIdx.x;
  const int Nvt = op.Nvt, Nx = op.Nx;
  const int r = s / Nvt, t = s - r * Nvt;
  const double av = op.axvr[r] * op.cos_vt[t];
  const bool fwd = av >= 0;
  const double* D = fwd ? op.x_l2r : op.x_r2l;
  const double dtav = dt * av;
  const long base = (long)s * Nx;
  double al[PER], be[PER];
  double A = 1.0, B = 0.0;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int p = lane * PER + i;
    al[i] = 0.0;
    be[i] = 1.0;                                 // identity beyond the end
    if (p < Nx) {
      const int c = fwd ? p : Nx - 1 - p;
      const double rden = 1.0 / (1.0 + dtav * D[c]);
      al[i] = p == 0 ? 0.0 : b[base + c] * rden;
      be[i] = p == 0 ? 0.0 : -dtav * D[Nx + (fwd ? c - 1 : c + 1)] * rden;
    }
    B = fma(be[i], B, al[i]);
    A *= be[i];
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double Ap = __shfl_up(A, off, 64), Bp = __shfl_up(B, off, 64);
    if (lane >= off) {
      B = fma(A, Bp, B);
      A *= Ap;
    }
  }
  double y = __shfl_up(B, 1, 64);
  if (lane == 0) y = 0.0;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int p = lane * PER + i;
    if (p < Nx) {
      y = fma(be[i], y, al[i]);
      out[base + (fwd ? p : Nx - 1 - p)] = y;
    }
  }
}

template <int MODE, int ND>
void launch_sweep_nd(const BzFvm1d3v& op, int N, int nsys, const double* E, double dt, const double* b, double* out,
                     const double* h, hipStream_t st) {
  const size_t lds = (size_t)(2 * ND + (MODE == SW_VR ? 1 : 0)) * N;
  const dim3 grid((nsys + SWT - 1) / SWT), block(SWT);
  if (lds <= SW_LDS_DOUBLES)
    hipLaunchKernelGGL((k_fvm_sweep<MODE, ND, true>), grid, block, lds * sizeof(double), st, op, E, dt, b, out, h);
  else
    hipLaunchKernelGGL((k_fvm_sweep<MODE, ND, false>), grid, block, 0, st, op, E, dt, b, out, h);
}

template <int MODE>
int launch_fvm_sweep(const BzFvm1d3v& op, const double* E, double dt, const double* b, double* out, const double* h,
                     hipStream_t st) {
  const int N = MODE == SW_VR ? op.Nr : (MODE == SW_VT ? op.Nvt : op.Nx);
  const int nsys = MODE == SW_VR ? op.Nvt * op.Nx : (MODE == SW_VT ? op.Nr * op.Nx : op.Nr * op.Nvt);
  const int nd = MODE == SW_VR ? op.nvr : (MODE == SW_VT ? op.nvt : op.nx);
  switch (nd) {
    case 1: launch_sweep_nd<MODE, 1>(op, N, nsys, E, dt, b, out, h, st); break;
    case 2: launch_sweep_nd<MODE, 2>(op, N, nsys, E, dt, b, out, h, st); break;
    case 3: launch_sweep_nd<MODE, 3>(op, N, nsys, E, dt, b, out, h, st); break;
    case 4: launch_sweep_nd<MODE, 4>(op, N, nsys, E, dt, b, out, h, st); break;
    case 5: launch_sweep_nd<MODE, 5>(op, N, nsys, E, dt, b, out, h, st); break;
    case 6: launch_sweep_nd<MODE, 6>(op, N, nsys, E, dt, b, out, h, st); break;
    case 7: launch_sweep_nd<MODE, 7>(op, N, nsys, E, dt, b, out, h, st); break;
    case 8: launch_sweep_nd<MODE, 8>(op, N, nsys, E, dt, b, out, h, st); break;
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
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  const int tx = op->Nx <= 64 ? 64 : (op->Nx <= 128 ? 128 : 256);       // row kernels: x per workgroup
  const long nwg = (long)op->Nr * op->Nvt * ((op->Nx + tx - 1) / tx);
  const dim3 growx((unsigned)((nwg + 7) / 8 * 8)), browx(tx);
  const bool rows_fit = nwg < (1L << 30);
  if (rows_fit && op->nx == 2 && op->nvr == 4 && op->nvt == 4)
    hipLaunchKernelGGL((k_fvm_matvec_w<2, 4, 4, 6>), growx, browx, 0, stream, *op, E, dt, x, s, y);
  else if (rows_fit && op->nx == 3 && op->nvr == 3 && op->nvt == 3)
    hipLaunchKernelGGL((k_fvm_matvec_w<3, 3, 3, 6>), growx, browx, 0, stream, *op, E, dt, x, s, y);
  else
    hipLaunchKernelGGL(k_fvm_matvec, grid, block, 0, stream, *op, E, dt, x, s, y);
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
  hipLaunchKernelGGL(k_fvm_coll_corr, dim3((op->Nr + 15) / 16, (op->Nx + 15) / 16), dim3(64 * CC_WAVES), 0, stream,
                     op->H, op->h_lo, s, op->Nr, op->Nx, h);
  BZ_LAUNCH_CHECK();
  // the radial sweep: segmented scan for the common stencils and long radial lines, else one
  // thread per line
  if (op->nvr >= 2 && op->nvr <= 5 && op->Nr >= 2 * NSEG)
    rc = launch_fvm_vr_scan(*op, E, alpha_dt, b, y, h, stream);
  else
    rc = launch_fvm_sweep<SW_VR>(*op, E, alpha_dt, b, y, h, stream);
  if (rc != BZ_OK) return rc;
  rc = launch_fvm_sweep<SW_VT>(*op, E, alpha_dt, y, y, nullptr, stream);
  if (rc != BZ_OK) return rc;
  if (op->nx == 2 && op->Nx <= 64 * 8) {
    const dim3 grid(op->Nr * op->Nvt), block(64);
    const int per = (op->Nx + 63) / 64;
    if (per <= 2) hipLaunchKernelGGL((k_fvm_x_scan<2>), grid, block, 0, stream, *op, alpha_dt, y, y);
    else if (per <= 4) hipLaunchKernelGGL((k_fvm_x_scan<4>), grid, block, 0, stream, *op, alpha_dt, y, y);
    else hipLaunchKernelGGL((k_fvm_x_scan<8>), grid, block, 0, stream, *op, alpha_dt, y, y);
    BZ_LAUNCH_CHECK();
    return BZ_OK;
  }
  return launch_fvm_sweep<SW_X>(*op, E, alpha_dt, y, y, nullptr, stream);
}

}  // extern "C"
