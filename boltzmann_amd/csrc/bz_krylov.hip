// Building blocks of restarted GMRES on the device (gfx950, FP64).
//
// The Krylov basis is stored row by row, V[j*ldv + e] (one contiguous vector per row),
// so every operation below streams whole vectors with consecutive threads on
// consecutive elements.  Reductions are two-pass and deterministic: each workgroup
// writes the partial sum of its chunk, and one wavefront per output adds the partials
// in a fixed order.  Orthogonalisation is classical Gram-Schmidt applied twice (two
// block dot products and two block updates per Arnoldi step, independent of the step
// count), which has the stability of modified Gram-Schmidt in FP64.
#include <algorithm>

#include "bz_common.h"

namespace {

constexpr int NT = 256;

// partial[j * nblk + blockIdx.x] = sum over this block's chunk of V[j][e] * u[e]; grid (nblk, nv)
__global__ __launch_bounds__(NT) void k_multi_dot_partial(const double* __restrict__ V, long ldv,
                                                          const double* __restrict__ u, long n, long chunk,
                                                          double* __restrict__ partial) {
  __shared__ double red[NT / 64];
  const int j = blockIdx.y;
  const long a = (long)blockIdx.x * chunk, b = std::min(n, a + chunk);
  const double* v = V + (long)j * ldv;
  double acc = 0.0;
  for (long e = a + threadIdx.x; e < b; e += NT) acc = fma(v[e], u[e], acc);
  acc = bz::wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < NT / 64; ++w) s += red[w];
    partial[(long)j * gridDim.x + blockIdx.x] = s;
  }
}

// out[j * ostride] = (accumulate ? out : 0) + s_j with s_j = sum_b partial[j * nblk + b] (or its
// square root); out2[j] = s_j as well when out2 is given
__global__ void k_reduce_partials(const double* __restrict__ partial, int nblk, int nv, double* out, int ostride,
                                  int accumulate, int take_sqrt, double* out2) {
  const int j = blockIdx.x;
  if (j >= nv) return;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64) acc += partial[(long)j * nblk + b];
  acc = bz::wave_sum(acc);
  if (threadIdx.x == 0) {
    const double v = take_sqrt ? sqrt(acc) : acc;
    if (out2) out2[j] = v;
    out[(long)j * ostride] = accumulate ? out[(long)j * ostride] + v : v;
  }
}

// dst[e] = src[e] + sign * sum_j coef[j * cstride] * V[j][e]
__global__ void k_vt_axpy(const double* __restrict__ V, long ldv, int nv, const double* __restrict__ coef, int cstride,
                          double sign, const double* src, double* dst, long n) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < nv; ++j) s = fma(coef[(long)j * cstride], V[(long)j * ldv + e], s);
    dst[e] = src[e] + sign * s;
  }
}

// y = x / s[0]; nothing is written when s[0] == 0 (Arnoldi breakdown)
__global__ void k_scale_inv(const double* __restrict__ x, const double* __restrict__ s, double* y, long n) {
  const double d = s[0];
  if (d == 0.0) return;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) y[e] = x[e] / d;
}

// min_y |H y - beta e_0| for the (m+1) x m upper Hessenberg H (row stride ldh), by Givens
// rotations on a copy; one thread (m is the restart length, at most a few dozen).
__global__ void k_hess_lstsq(const double* __restrict__ H, int ldh, int m, const double* __restrict__ beta,
                             double* __restrict__ R, double* __restrict__ y) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  // R: (m+1) x m work copy, g: right-hand side stored in y[m+1 ...]
  double* g = y + m;
  for (int r = 0; r <= m; ++r)
    for (int c = 0; c < m; ++c) R[r * m + c] = H[(long)r * ldh + c];
  for (int r = 0; r <= m; ++r) g[r] = 0.0;
  g[0] = beta[0];
  for (int c = 0; c < m; ++c) {
    const double a = R[c * m + c], b = R[(c + 1) * m + c];
    const double rr = hypot(a, b);
    const double cs = (rr == 0.0) ? 1.0 : a / rr, sn = (rr == 0.0) ? 0.0 : b / rr;
    for (int k = c; k < m; ++k) {
      const double t1 = R[c * m + k], t2 = R[(c + 1) * m + k];
      R[c * m + k] = cs * t1 + sn * t2;
      R[(c + 1) * m + k] = -sn * t1 + cs * t2;
    }
    const double g1 = g[c], g2 = g[c + 1];
    g[c] = cs * g1 + sn * g2;
    g[c + 1] = -sn * g1 + cs * g2;
  }
  for (int c = m - 1; c >= 0; --c) {
    double s = g[c];
    for (int k = c + 1; k < m; ++k) s -= R[c * m + k] * y[k];
    y[c] = (R[c * m + c] != 0.0) ? s / R[c * m + c] : 0.0;
  }
}

int grid_for(long n) { return (int)std::min<long>((n + NT - 1) / NT, 8192); }

// bin[x] = argmin_i |E[x] - Ev[i]| (first minimum), the preconditioner membership of
// glow1d_boltzmann.vspace_pc_setup
__global__ void k_nearest_bin(const double* __restrict__ E, int nx, const double* __restrict__ Ev, int nb,
                              int* __restrict__ bin) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= nx) return;
  const double e = E[x];
  int best = 0;
  double bd = fabs(e - Ev[0]);
  for (int i = 1; i < nb; ++i) {
    const double d = fabs(e - Ev[i]);
    if (d < bd) { bd = d; best = i; }
  }
  bin[x] = best;
}

// y[i, x] = sum_k P[bin[x]][i, k] u[k, x]; u, y (dof, nx) with x fastest.  Consecutive
// threads take consecutive x; neighbouring points share a bin, so the rows of P they
// read are the same cache lines.
__global__ void k_binned_block_apply(const double* __restrict__ P, int dof, const int* __restrict__ bin,
                                     const double* __restrict__ u, int nx, double* __restrict__ y) {
  const long total = (long)dof * nx;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int x = (int)(e % nx), i = (int)(e / nx);
    const double* row = P + ((long)bin[x] * dof + i) * dof;
    double s = 0.0;
    for (int k = 0; k < dof; ++k) s = fma(row[k], u[(long)k * nx + x], s);
    y[e] = s;
  }
}

// y = x - c (Kx + E[col] Ax) for (dof, nx) arrays: the backward-Euler v-space operator
// I - dt tau (n0 np0 (C + Tg C_T) + E A) applied from the two operator products
__global__ void k_vstep_combine(const double* __restrict__ x, const double* __restrict__ Kx,
                                const double* __restrict__ Ax, const double* __restrict__ E, double c, int nx,
                                long total, double* __restrict__ y) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
    y[e] = x[e] - c * (Kx[e] + E[e % nx] * Ax[e]);
}

// ---------------------------------------------------------------- batched GMRES (many small systems)
// B independent systems of size n, one wavefront each.  Basis V[(j * B + b) * n + e],
// j = 0..m; the Hessenberg columns are reduced by Givens rotations as they arrive, so
// R[b][col * m + row] holds the triangular factor, g[b][0..m] the rotated right-hand
// side and |g[b][k]| the residual norm after k steps.

__device__ __forceinline__ double wdot(const double* __restrict__ a, const double* __restrict__ b, int n) {
  double s = 0.0;
  for (int e = threadIdx.x; e < n; e += 64) s = fma(a[e], b[e], s);
  return bz::wave_sum(s);
}

// Arnoldi step i of every active system: orthogonalise w[b] (the operator applied to
// the preconditioned V_i) against V_0..V_i (classical Gram-Schmidt, twice), append
// V_{i+1}, rotate the new column, update the residual estimate; a system whose residual
// reaches target[b] (or whose basis breaks down) becomes inactive with kdim = i + 1.
__global__ __launch_bounds__(64) void k_arnoldi_batched(int n, int B, int m, int i, double* __restrict__ V,
                                                        const double* __restrict__ w, double* __restrict__ R,
                                                        double* __restrict__ cs, double* __restrict__ sn,
                                                        double* __restrict__ g, const double* __restrict__ target,
                                                        int* __restrict__ active, int* __restrict__ kdim,
                                                        double* __restrict__ resid) {
  extern __shared__ double lds[];
  const int inst = blockIdx.x, lane = threadIdx.x;
  if (!active[inst]) return;
  double* W = lds;
  double* HC = W + n;          // [m + 2] Hessenberg column
  double* T = HC + m + 2;      // [m + 1] projections of one pass
  for (int e = lane; e < n; e += 64) W[e] = w[(size_t)inst * n + e];
  for (int j = lane; j <= i + 1; j += 64) HC[j] = 0.0;
  __syncthreads();
  for (int rep = 0; rep < 2; ++rep) {
    for (int j = 0; j <= i; ++j) {
      const double d = wdot(V + ((size_t)j * B + inst) * n, W, n);
      if (lane == 0) { T[j] = d; HC[j] += d; }
    }
    __syncthreads();
    for (int e = lane; e < n; e += 64) {
      double s = W[e];
      for (int j = 0; j <= i; ++j) s = fma(-T[j], V[((size_t)j * B + inst) * n + e], s);
      W[e] = s;
    }
    __syncthreads();
  }
  const double hn = sqrt(wdot(W, W, n));
  const double inv = hn > 0.0 ? 1.0 / hn : 0.0;
  double* Vn = V + ((size_t)(i + 1) * B + inst) * n;
  for (int e = lane; e < n; e += 64) Vn[e] = W[e] * inv;
  if (lane == 0) {
    HC[i + 1] = hn;
    double* csb = cs + (size_t)inst * m;
    double* snb = sn + (size_t)inst * m;
    double* gb = g + (size_t)inst * (m + 1);
    for (int j = 0; j < i; ++j) {
      const double a = HC[j], b2 = HC[j + 1];
      HC[j] = fma(csb[j], a, snb[j] * b2);
      HC[j + 1] = fma(-snb[j], a, csb[j] * b2);
    }
    const double a = HC[i], b2 = HC[i + 1];
    const double r = hypot(a, b2);
    const double c = r > 0.0 ? a / r : 1.0, s = r > 0.0 ? b2 / r : 0.0;
    csb[i] = c;
    snb[i] = s;
    HC[i] = r;
    const double gi = gb[i];
    gb[i] = c * gi;
    gb[i + 1] = -s * gi;
    double* Rb = R + (size_t)inst * m * m + (size_t)i * m;
    for (int j = 0; j <= i; ++j) Rb[j] = HC[j];
    const double res = fabs(gb[i + 1]);
    resid[inst] = res;
    kdim[inst] = i + 1;
    if (res <= target[inst] || !(hn > 0.0) || i + 1 >= m) active[inst] = 0;
  }
}

// out[b] = sum_j y_j V_j[b] with R y = g (the first kdim[b] steps); zero when kdim[b] = 0
__global__ __launch_bounds__(64) void k_gmres_combine_batched(int n, int B, int m, const double* __restrict__ V,
                                                              const double* __restrict__ R,
                                                              const double* __restrict__ g,
                                                              const int* __restrict__ kdim, double* __restrict__ out) {
  __shared__ double Y[64];
  const int inst = blockIdx.x, lane = threadIdx.x;
  const int k = kdim[inst];
  if (lane == 0) {
    const double* Rb = R + (size_t)inst * m * m;
    const double* gb = g + (size_t)inst * (m + 1);
    for (int r = k - 1; r >= 0; --r) {
      double s = gb[r];
      for (int c = r + 1; c < k; ++c) s = fma(-Rb[(size_t)c * m + r], Y[c], s);
      Y[r] = Rb[(size_t)r * m + r] != 0.0 ? s / Rb[(size_t)r * m + r] : 0.0;
    }
  }
  __syncthreads();
  for (int e = lane; e < n; e += 64) {
    double s = 0.0;
    for (int j = 0; j < k; ++j) s = fma(Y[j], V[((size_t)j * B + inst) * n + e], s);
    out[(size_t)inst * n + e] = s;
  }
}

}  // namespace

extern "C" {

int bz_multi_dot(const double* V, long ldv, int nv, const double* u, long n, double* partial, int nblk, double* out,
                 int ostride, int accumulate, int take_sqrt, double* out2, hipStream_t stream) {
  if (nv < 1 || n < 1 || nblk < 1 || ldv < n || !V || !u || !partial || !out || ostride < 1) return BZ_EINVAL;
  const long chunk = (n + nblk - 1) / nblk;
  hipLaunchKernelGGL(k_multi_dot_partial, dim3(nblk, nv), dim3(NT), 0, stream, V, ldv, u, n, chunk, partial);
  BZ_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_partials, dim3(nv), dim3(64), 0, stream, partial, nblk, nv, out, ostride, accumulate,
                     take_sqrt, out2);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_vt_axpy(const double* V, long ldv, int nv, const double* coef, int cstride, double sign, const double* src,
               double* dst, long n, hipStream_t stream) {
  if (nv < 0 || n < 1 || ldv < n || !V || !coef || !src || !dst || cstride < 1) return BZ_EINVAL;
  hipLaunchKernelGGL(k_vt_axpy, dim3(grid_for(n)), dim3(NT), 0, stream, V, ldv, nv, coef, cstride, sign, src, dst, n);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_scale_inv(const double* x, const double* s, double* y, long n, hipStream_t stream) {
  if (n < 1 || !x || !s || !y) return BZ_EINVAL;
  hipLaunchKernelGGL(k_scale_inv, dim3(grid_for(n)), dim3(NT), 0, stream, x, s, y, n);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_nearest_bin(const double* E, int nx, const double* Ev, int nb, int* bin, hipStream_t stream) {
  if (nx < 0 || nb < 1 || !E || !Ev || !bin) return BZ_EINVAL;
  if (nx == 0) return BZ_OK;
  hipLaunchKernelGGL(k_nearest_bin, dim3((nx + 255) / 256), dim3(256), 0, stream, E, nx, Ev, nb, bin);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_binned_block_apply(const double* P, int dof, const int* bin, const double* u, int nx, double* y,
                          hipStream_t stream) {
  if (dof < 1 || nx < 1 || !P || !bin || !u || !y) return BZ_EINVAL;
  hipLaunchKernelGGL(k_binned_block_apply, dim3(grid_for((long)dof * nx)), dim3(NT), 0, stream, P, dof, bin, u, nx, y);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_vstep_combine(const double* x, const double* Kx, const double* Ax, const double* E, double c, int dof, int nx,
                     double* y, hipStream_t stream) {
  if (dof < 1 || nx < 1 || !x || !Kx || !Ax || !E || !y) return BZ_EINVAL;
  const long total = (long)dof * nx;
  hipLaunchKernelGGL(k_vstep_combine, dim3(grid_for(total)), dim3(NT), 0, stream, x, Kx, Ax, E, c, nx, total, y);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_hess_lstsq(const double* H, int ldh, int m, const double* beta, double* work, double* y, hipStream_t stream) {
  if (m < 1 || ldh < m || !H || !beta || !work || !y) return BZ_EINVAL;
  hipLaunchKernelGGL(k_hess_lstsq, dim3(1), dim3(64), 0, stream, H, ldh, m, beta, work, y);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_arnoldi_batched(int n, int B, int m, int i, double* V, const double* w, double* R, double* cs, double* sn,
                       double* g, const double* target, int* active, int* kdim, double* resid, hipStream_t stream) {
  if (n < 1 || B < 1 || m < 1 || m > 64 || i < 0 || i >= m || !V || !w || !R || !cs || !sn || !g || !target ||
      !active || !kdim || !resid)
    return BZ_EINVAL;
  const size_t lds = (size_t)(n + 2 * m + 3) * sizeof(double);
  if (lds > 64 * 1024) return BZ_EINVAL;
  hipLaunchKernelGGL(k_arnoldi_batched, dim3(B), dim3(64), lds, stream, n, B, m, i, V, w, R, cs, sn, g, target, active,
                     kdim, resid);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_gmres_combine_batched(int n, int B, int m, const double* V, const double* R, const double* g, const int* kdim,
                             double* out, hipStream_t stream) {
  if (n < 1 || B < 1 || m < 1 || m > 64 || !V || !R || !g || !kdim || !out) return BZ_EINVAL;
  hipLaunchKernelGGL(k_gmres_combine_batched, dim3(B), dim3(64), 0, stream, n, B, m, V, R, g, kdim, out);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

}  // extern "C"
