// Galerkin assembly of the B-spline (|v|) x spherical-harmonic operators, FP64.
//
// One thread per (test p, trial k) radial pair; each thread walks the Gauss
// points of the <= 4 knot intervals of the shared support, evaluating the 4
// non-zero cubic B-splines by the de Boor recursion.  All radial blocks are
// written dense (np_r x np_r, row-major, row = test function p).
//
// Reference algorithms (BESolver/python):
//   mass / advection radial integrals   spec_spherical.py:226-277, 442-498
//   binary collision operator (CG)      collision_operator_spherical.py:96-223
//   gas-temperature operator (CG)       collision_operator_spherical.py:443-497
//   QoI vectors                         utils.py:73-150, 321-395, 854-944
//   Cholesky inverse of the mass matrix spec_spherical.py:299-336
#include "bz_common.h"
#include "../../include/bz.h"

namespace {

using namespace bz;

// Points of interval i (non-degenerate [t[i], t[i+1]]): x_j, w_j from reference nodes.
__device__ __forceinline__ void quad_point(const double* __restrict__ t, int i, const double* __restrict__ glx,
                                           const double* __restrict__ glw, int j, double& x, double& w) {
  const double a = t[i], b = t[i + 1];
  x = 0.5 * (b - a) * glx[j] + 0.5 * (b + a);
  w = 0.5 * (b - a) * glw[j];
}

// out[p*npr + k] = sum_x w x^pw B_p^(da)(x) B_k^(db)(x)  over supp(B_p) cap supp(B_k)
__global__ void k_radial_bilinear(const double* __restrict__ t, int npr, const double* __restrict__ glx,
                                  const double* __restrict__ glw, int q, int pw, int da, int db,
                                  double* __restrict__ out) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= npr * npr) return;
  const int p = gid / npr, k = gid % npr;
  double acc = 0.0;
  const int lo = max(max(p, k), 3);
  const int hi = min(min(p, k) + 3, npr - 1);
  for (int i = lo; i <= hi; ++i) {
    if (!(t[i] < t[i + 1])) continue;
    for (int j = 0; j < q; ++j) {
      double x, w;
      quad_point(t, i, glx, glw, j, x, w);
      double N[4], D[4];
      bspline3_eval(t, i, x, N, D);
      const double bp = da ? D[p - (i - 3)] : N[p - (i - 3)];
      const double bk = db ? D[k - (i - 3)] : N[k - (i - 3)];
      double xp = 1.0;
      for (int e = 0; e < pw; ++e) xp *= x;
      acc += w * xp * bp * bk;
    }
  }
  out[gid] = acc;
}

// Binary electron-heavy collision radial blocks (q0: l = 0, q1: l >= 1).
// ptype: 0 elastic, 1 excitation, 2 ionization, 3 attachment.
__global__ void k_collision_radial(const double* __restrict__ t, int npr, const double* __restrict__ glx,
                                   const double* __restrict__ glw, int q, int ptype, double vth, double c_gamma,
                                   double mByM, double threshold, double att_energy, double kappa,
                                   const double* __restrict__ tab_e, const double* __restrict__ tab_s, int tab_len,
                                   double* __restrict__ q0, double* __restrict__ q1) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= npr * npr) return;
  const int p = gid / npr, k = gid % npr;
  const double split = (ptype == 2) ? 2.0 : 1.0;
  const double vscale = sqrt(1.0 - 2.0 * mByM);
  double a0 = 0.0, a1 = 0.0;
  const int lo = max(k, 3), hi = min(k + 3, npr - 1);   // intervals of supp(B_k)
  for (int i = lo; i <= hi; ++i) {
    if (!(t[i] < t[i + 1])) continue;
    for (int j = 0; j < q; ++j) {
      double x, w;
      quad_point(t, i, glx, glw, j, x, w);
      const double ev = (x * vth / c_gamma) * (x * vth / c_gamma);
      double vpost;
      if (ptype == 0) {
        vpost = x * vscale;
      } else if (ptype == 1 || ptype == 2) {
        if (!(ev >= threshold)) continue;     // sub-threshold points are dropped
        vpost = c_gamma * sqrt((1.0 / split) * (ev - threshold)) / vth;
      } else {
        vpost = c_gamma * sqrt((1.0 / split) * (ev + att_energy)) / vth;
      }
      const double sig = interp_linear_flat(tab_e, tab_s, tab_len, ev);
      double N[4], D[4];
      bspline3_eval(t, i, x, N, D);
      const int rk = k - (i - 3);
      const int rp = p - (i - 3);
      const double bk = N[rk];
      const double bp = (rp >= 0 && rp <= 3) ? N[rp] : 0.0;
      const double bpp = bspline3_single(t, npr, p, vpost, 0);
      const double x3 = x * x * x;
      const double wt = w * vth * x3 * sig;
      if (ptype == 0) {
        // B_p(v') - B_p(v) with v - v' = v (2m/M) / (1 + sqrt(1 - 2m/M)), free of cancellation
        const double dv = x * (2.0 * mByM) / (1.0 + vscale);
        const double db = bspline3_diff_down(t, npr, p, x, dv);
        double dterm = 0.0;
        if (kappa != 0.0 && rp >= 0 && rp <= 3) dterm = kappa * w * x3 * sig * D[rp] * D[rk];
        a0 += wt * bk * db - dterm;
      } else {
        a0 += wt * bk * (split * bpp - bp);
      }
      a1 += -wt * bk * bp;
    }
  }
  q0[gid] = a0;
  q1[gid] = a1;
}

// Gas-temperature operator radial block (l = 0 only), tgK = 1.
__global__ void k_gas_temperature_radial(const double* __restrict__ t, int npr, const double* __restrict__ glx,
                                         const double* __restrict__ glw, int q, double vth, double c_gamma,
                                         double kappa1, const double* __restrict__ tab_e,
                                         const double* __restrict__ tab_s, int tab_len, double* __restrict__ out) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= npr * npr) return;
  const int p = gid / npr, k = gid % npr;
  double acc = 0.0;
  const int lo = max(max(p, k), 3);
  const int hi = min(min(p, k) + 3, npr - 1);
  for (int i = lo; i <= hi; ++i) {
    if (!(t[i] < t[i + 1])) continue;
    for (int j = 0; j < q; ++j) {
      double x, w;
      quad_point(t, i, glx, glw, j, x, w);
      const double ev = (x * vth / c_gamma) * (x * vth / c_gamma);
      const double sig = interp_linear_flat(tab_e, tab_s, tab_len, ev);
      double N[4], D[4];
      bspline3_eval(t, i, x, N, D);
      acc += -kappa1 * w * x * x * x * sig * D[p - (i - 3)] * D[k - (i - 3)];
    }
  }
  out[gid] = acc;
}

// Radial moment vector: out[p] = scale * sum_x w x^pw g(x) B_p(x),
//   g = 1 (mode 0), sigma(x) (mode 1), 1/sum_m sigma_m(x) (mode 2, tables concatenated),
//   exp(-x^2) (mode 3, Maxwellian projection of function_to_basis, utils.py:475-532).
__global__ void k_radial_moment(const double* __restrict__ t, int npr, const double* __restrict__ glx,
                                const double* __restrict__ glw, int q, int pw, int mode, double scale,
                                double vth, double c_gamma, const double* __restrict__ tab_e,
                                const double* __restrict__ tab_s, const int* __restrict__ tab_off, int ntab,
                                double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npr) return;
  double acc = 0.0;
  const int lo = max(p, 3), hi = min(p + 3, npr - 1);
  for (int i = lo; i <= hi; ++i) {
    if (!(t[i] < t[i + 1])) continue;
    for (int j = 0; j < q; ++j) {
      double x, w;
      quad_point(t, i, glx, glw, j, x, w);
      double N[4], D[4];
      bspline3_eval(t, i, x, N, D);
      double g = 1.0;
      if (mode == 3) {
        g = exp(-x * x);
      } else if (mode != 0) {
        const double ev = (x * vth / c_gamma) * (x * vth / c_gamma);
        double s = 0.0;
        for (int m = 0; m < ntab; ++m) {
          const int o = tab_off[m], len = tab_off[m + 1] - tab_off[m];
          s += interp_linear_flat(tab_e + o, tab_s + o, len, ev);
        }
        g = (mode == 1) ? s : 1.0 / s;
      }
      double xp = 1.0;
      for (int e = 0; e < pw; ++e) xp *= x;
      acc += w * xp * g * N[p - (i - 3)];
    }
  }
  out[p] = scale * acc;
}

// Scatter radial blocks into the dense (n x n) operator, radial-major index p*ns + l.
// blk: nblk blocks of npr*npr; lmap[l] = block used on the (l,l) diagonal (-1: zero).
__global__ void k_scatter_diag_blocks(const double* __restrict__ blk, int npr, int ns,
                                      const int* __restrict__ lmap, double scale, double* __restrict__ out) {
  const int n = npr * ns;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)n * n) return;
  const int row = gid / n, col = gid % n;
  const int p = row / ns, l = row % ns, k = col / ns, l2 = col % ns;
  double v = 0.0;
  if (l == l2) {
    const int b = lmap[l];
    if (b >= 0) v = scale * blk[(long)b * npr * npr + p * npr + k];
  }
  out[gid] = v;
}

// E-field advection operator: (p,qi) <- (k, qi+1): mm1*AM + AD*mm2 ; (k, qi-1): mm1*BM + BD*mm2.
// cup[qi], cdn[qi]: (coefficient of mm1, coefficient of mm2) for the l = qi+1 and l = qi-1 couplings.
__global__ void k_advection_scatter(const double* __restrict__ mm1, const double* __restrict__ mm2, int npr, int ns,
                                    const double* __restrict__ cup, const double* __restrict__ cdn,
                                    double scale, double* __restrict__ out) {
  const int n = npr * ns;
  const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long)n * n) return;
  const int row = gid / n, col = gid % n;
  const int p = row / ns, qi = row % ns, k = col / ns, l = col % ns;
  double v = 0.0;
  const int e = p * npr + k;
  if (l == qi + 1) v = mm1[e] * cup[2 * qi] + cup[2 * qi + 1] * mm2[e];
  else if (l == qi - 1) v = mm1[e] * cdn[2 * qi] + cdn[2 * qi + 1] * mm2[e];
  out[gid] = scale * v;
}

// Banded Cholesky of one SPD radial block (half bandwidth hb), single thread per block;
// L stored as lb[p*(hb+1) + d] = L[p, p-d].
__global__ void k_band_cholesky(const double* __restrict__ A, int npr, int hb, double* __restrict__ lb) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int j = 0; j < npr; ++j) {
    for (int i = j; i <= min(npr - 1, j + hb); ++i) {
      double s = A[i * npr + j];
      for (int k = max(0, i - hb); k < j; ++k) s -= lb[i * (hb + 1) + (i - k)] * lb[j * (hb + 1) + (j - k)];
      if (i == j) lb[j * (hb + 1)] = sqrt(s);
      else lb[i * (hb + 1) + (i - j)] = s / lb[j * (hb + 1)];
    }
  }
}

// Column c of the inverse of L L^T (banded L from k_band_cholesky), one thread per column.
__global__ void k_band_chol_inverse(const double* __restrict__ lb, int npr, int hb, double* __restrict__ inv,
                                    double* __restrict__ work) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= npr) return;
  double* y = work + (long)c * npr;
  for (int i = 0; i < npr; ++i) {
    double s = (i == c) ? 1.0 : 0.0;
    for (int k = max(0, i - hb); k < i; ++k) s -= lb[i * (hb + 1) + (i - k)] * y[k];
    y[i] = s / lb[i * (hb + 1)];
  }
  for (int i = npr - 1; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k <= min(npr - 1, i + hb); ++k) s -= lb[k * (hb + 1) + (k - i)] * y[k];
    y[i] = s / lb[i * (hb + 1)];
  }
  for (int i = 0; i < npr; ++i) inv[i * npr + c] = y[i];
}

// Solve (L L^T) y = b for one column, L banded from k_band_cholesky (hb sub-diagonals).
__device__ void band_chol_solve_col(const double* __restrict__ lb, int npr, int hb, double* __restrict__ y) {
  for (int i = 0; i < npr; ++i) {
    double s = y[i];
    for (int k = max(0, i - hb); k < i; ++k) s -= lb[i * (hb + 1) + (i - k)] * y[k];
    y[i] = s / lb[i * (hb + 1)];
  }
  for (int i = npr - 1; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k <= min(npr - 1, i + hb); ++k) s -= lb[k * (hb + 1) + (k - i)] * y[k];
    y[i] = s / lb[i * (hb + 1)];
  }
}

// out[:, c] = A^-1 rhs[:, c] for the banded SPD A (dense storage, |i - k| <= hb), one
// thread per column: a Cholesky solve, then `refine` steps of iterative refinement with
// the residual rhs - A y summed in compensated arithmetic (two_sum_acc).  The refined
// result is the solution for the FP64 A and rhs to about one rounding, where forming
// A^-1 explicitly and multiplying loses cond(A) roundings (the B-spline mass matrix with
// its v^2 weight is ill-conditioned by its scaling near v = 0).
__global__ void k_band_chol_solve_refined(const double* __restrict__ A, const double* __restrict__ lb, int npr,
                                          int hb, const double* __restrict__ rhs, int ncol, double* __restrict__ out,
                                          double* __restrict__ work, int refine) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  double* y = work + (long)c * 2 * npr;
  double* r = y + npr;
  for (int i = 0; i < npr; ++i) y[i] = rhs[(long)i * ncol + c];
  band_chol_solve_col(lb, npr, hb, y);
  for (int it = 0; it < refine; ++it) {
    for (int i = 0; i < npr; ++i) {
      double s = rhs[(long)i * ncol + c], e = 0.0;
      for (int k = max(0, i - hb); k <= min(npr - 1, i + hb); ++k) two_sum_acc(-A[(long)i * npr + k], y[k], s, e);
      r[i] = s + e;
    }
    band_chol_solve_col(lb, npr, hb, r);
    for (int i = 0; i < npr; ++i) y[i] += r[i];
  }
  for (int i = 0; i < npr; ++i) out[(long)i * ncol + c] = y[i];
}

// out[k*nx + i] = B_k^(dorder)(x_i): basis Vandermonde (spec_spherical.py:338-383, Vq_r / Vdq_r).
__global__ void k_basis_table(const double* __restrict__ t, int npr, const double* __restrict__ x, int nx,
                              int dorder, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  for (int k = 0; k < npr; ++k) out[(long)k * nx + i] = 0.0;
  const int l = find_interval(t, npr, 3, x[i]);
  if (l < 0) return;
  double N[4], D[4];
  bspline3_eval(t, l, x[i], N, D);
  for (int r = 0; r < 4; ++r) out[(long)(l - 3 + r) * nx + i] = dorder ? D[r] : N[r];
}

inline int nblocks(long n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

extern "C" {

int bz_radial_bilinear(const double* t, int npr, const double* glx, const double* glw, int q, int pw, int da,
                       int db, double* out, hipStream_t stream) {
  if (npr < 4 || q < 1 || pw < 0 || (da != 0 && da != 1) || (db != 0 && db != 1)) return BZ_EINVAL;
  hipLaunchKernelGGL(k_radial_bilinear, dim3(nblocks((long)npr * npr, 256)), dim3(256), 0, stream, t, npr, glx,
                     glw, q, pw, da, db, out);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_collision_radial(const double* t, int npr, const double* glx, const double* glw, int q, int ptype,
                        double vth, double c_gamma, double mByM, double threshold, double att_energy, double kappa,
                        const double* tab_e, const double* tab_s, int tab_len, double* q0, double* q1,
                        hipStream_t stream) {
  if (npr < 4 || q < 1 || ptype < 0 || ptype > 3 || tab_len < 2) return BZ_EINVAL;
  hipLaunchKernelGGL(k_collision_radial, dim3(nblocks((long)npr * npr, 256)), dim3(256), 0, stream, t, npr, glx,
                     glw, q, ptype, vth, c_gamma, mByM, threshold, att_energy, kappa, tab_e, tab_s, tab_len, q0, q1);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_gas_temperature_radial(const double* t, int npr, const double* glx, const double* glw, int q, double vth,
                              double c_gamma, double kappa1, const double* tab_e, const double* tab_s, int tab_len,
                              double* out, hipStream_t stream) {
  if (npr < 4 || q < 1 || tab_len < 2) return BZ_EINVAL;
  hipLaunchKernelGGL(k_gas_temperature_radial, dim3(nblocks((long)npr * npr, 256)), dim3(256), 0, stream, t, npr,
                     glx, glw, q, vth, c_gamma, kappa1, tab_e, tab_s, tab_len, out);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_radial_moment(const double* t, int npr, const double* glx, const double* glw, int q, int pw, int mode,
                     double scale, double vth, double c_gamma, const double* tab_e, const double* tab_s,
                     const int* tab_off, int ntab, double* out, hipStream_t stream) {
  if (npr < 4 || q < 1 || pw < 0 || mode < 0 || mode > 3 || ((mode == 1 || mode == 2) && ntab < 1)) return BZ_EINVAL;
  hipLaunchKernelGGL(k_radial_moment, dim3(nblocks(npr, 128)), dim3(128), 0, stream, t, npr, glx, glw, q, pw, mode,
                     scale, vth, c_gamma, tab_e, tab_s, tab_off, ntab, out);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_scatter_diag_blocks(const double* blk, int npr, int ns, const int* lmap, double scale, double* out,
                           hipStream_t stream) {
  if (npr < 1 || ns < 1) return BZ_EINVAL;
  const long n = (long)npr * ns;
  hipLaunchKernelGGL(k_scatter_diag_blocks, dim3(nblocks(n * n, 256)), dim3(256), 0, stream, blk, npr, ns, lmap,
                     scale, out);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_advection_scatter(const double* mm1, const double* mm2, int npr, int ns, const double* cup,
                         const double* cdn, double scale, double* out, hipStream_t stream) {
  if (npr < 1 || ns < 1) return BZ_EINVAL;
  const long n = (long)npr * ns;
  hipLaunchKernelGGL(k_advection_scatter, dim3(nblocks(n * n, 256)), dim3(256), 0, stream, mm1, mm2, npr, ns, cup,
                     cdn, scale, out);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_basis_table(const double* t, int npr, const double* x, int nx, int dorder, double* out, hipStream_t stream) {
  if (npr < 4 || nx < 0 || (dorder != 0 && dorder != 1)) return BZ_EINVAL;
  if (nx == 0) return BZ_OK;
  hipLaunchKernelGGL(k_basis_table, dim3(nblocks(nx, 128)), dim3(128), 0, stream, t, npr, x, nx, dorder, out);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_band_cholesky(const double* A, int npr, int hb, double* lb, hipStream_t stream) {
  if (npr < 1 || hb < 0) return BZ_EINVAL;
  hipLaunchKernelGGL(k_band_cholesky, dim3(1), dim3(64), 0, stream, A, npr, hb, lb);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_band_chol_solve_refined(const double* A, const double* lb, int npr, int hb, const double* rhs, int ncol,
                               double* out, double* work, int refine, hipStream_t stream) {
  if (npr < 1 || hb < 0 || ncol < 1 || refine < 0) return BZ_EINVAL;
  hipLaunchKernelGGL(k_band_chol_solve_refined, dim3(nblocks(ncol, 64)), dim3(64), 0, stream, A, lb, npr, hb, rhs,
                     ncol, out, work, refine);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_band_chol_inverse(const double* lb, int npr, int hb, double* inv, double* work, hipStream_t stream) {
  if (npr < 1 || hb < 0) return BZ_EINVAL;
  hipLaunchKernelGGL(k_band_chol_inverse, dim3(nblocks(npr, 64)), dim3(64), 0, stream, lb, npr, hb, inv, work);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

}  // extern "C"
