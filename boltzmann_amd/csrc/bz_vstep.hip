// The backward-Euler v-step of the 1D3V glow step (step_bte_v, glowdischarge_boltzmann_1d.py:
// 1476-1592) as device kernels for a graph-captured GMRES cycle (gfx950, FP64).
//
//   matvec      Y = Z - c (K0 Z + E_x (A Z))          two (dof x dof) x (dof x Nx) products
//   precond     Y[:, x] = P_bin(x) Z[:, x]            one (dof x dof) block per E bin
//   givens      one Givens rotation per Arnoldi step, the residual estimate and a stop flag
//
// The products use the FP64 matrix cores (v_mfma_f64_16x16x4f64), one workgroup per 16 x 16
// output tile whose four wavefronts each take a quarter of K (summed through LDS), operands
// loaded straight from global memory with eight K-steps of loads in flight (the 387 x 387
// operators stay in L2).  MFMA operand layout (checked on the
// device): A[i][k] from lane 16 k + i, B[k][j] from lane 16 k + j, and the four accumulator
// values of lane l hold D[4 r + l / 16][l % 16], r = 0..3.
// The preconditioner groups the x points by bin (k_group_bins, once per v-step); its grid
// has a fixed shape (row blocks x column blocks x bins) so that it can be captured once, and
// blocks past a bin's point count exit at once.  Every kernel of the Arnoldi step reads a
// device flag and does nothing once the cycle has converged, so a captured step may be
// replayed past convergence.
#include "bz_common.h"
#include "../../include/bz.h"

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int UK = 8;       // K steps (of 4) whose operand loads are issued together

// One wavefront computes a 16 x 16 tile of OP Z for NOP operators (row-major dof x dof)
// and the Z columns cols[0..15] (gathered), loading the MFMA operands straight from global
// memory (the operators stay in L2 between launches): A[i][k] = OP[row0 + i][k] from lane
// 16 k + i, B[k][j] = Z[k][cols[j]] from lane 16 k + j.  UK steps of loads are in flight
// before their UK MFMAs.
template <int NOP>
__device__ __forceinline__ void tile_product(const double* const* ops, int dof, const double* __restrict__ Z,
                                             int ldz, int col, bool colok, int row0, int kbeg, int kend,
                                             v4d acc[NOP]) {
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int row = row0 + li;
  const bool rowok = row < dof;
  const size_t rbase = (size_t)(rowok ? row : 0) * dof;
  for (int k0 = kbeg; k0 < kend; k0 += 4 * UK) {
    double a[NOP][UK], b[UK];
#pragma unroll
    for (int u = 0; u < UK; ++u) {
      const int k = k0 + 4 * u + lk;
      const bool kok = k < kend;
      const int kc = kok ? k : 0;
#pragma unroll
      for (int o = 0; o < NOP; ++o) a[o][u] = ops[o][rbase + kc];
      b[u] = Z[(size_t)kc * ldz + col];
      if (!(kok && rowok)) {
#pragma unroll
        for (int o = 0; o < NOP; ++o) a[o][u] = 0.0;
      }
      if (!(kok && colok)) b[u] = 0.0;
    }
#pragma unroll
    for (int u = 0; u < UK; ++u)
#pragma unroll
      for (int o = 0; o < NOP; ++o) acc[o] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[o][u], b[u], acc[o], 0, 0, 0);
  }
}

// The four wavefronts of a workgroup take a quarter of K each (in whole UK groups); their
// partial tiles are summed through LDS into wavefront 0.  Returns true on wavefront 0.
constexpr int NSPLIT = 4;

template <int NOP>
__device__ __forceinline__ bool split_k_product(const double* const* ops, int dof, const double* __restrict__ Z,
                                                int ldz, int col, bool colok, int row0, v4d acc[NOP],
                                                double* red) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int groups = (dof + 4 * UK - 1) / (4 * UK);
  const int per = (groups + NSPLIT - 1) / NSPLIT;
  const int kbeg = min(dof, w * per * 4 * UK), kend = min(dof, (w + 1) * per * 4 * UK);
  tile_product<NOP>(ops, dof, Z, ldz, col, colok, row0, kbeg, kend, acc);
  if (w > 0) {
#pragma unroll
    for (int o = 0; o < NOP; ++o)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[(((w - 1) * NOP + o) * 4 + r) * 64 + lane] = acc[o][r];
  }
  __syncthreads();
  if (w > 0) return false;
#pragma unroll
  for (int q = 1; q < NSPLIT; ++q)
#pragma unroll
    for (int o = 0; o < NOP; ++o)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[o][r] += red[(((q - 1) * NOP + o) * 4 + r) * 64 + lane];
  return true;
}

// Y = Z - c (K0 Z + E_x (A Z)), Z and Y (dof x nx) row-major; one workgroup per 16 x 16 tile
__global__ __launch_bounds__(64 * NSPLIT) void k_vstep_matvec(const double* __restrict__ K0,
                                                              const double* __restrict__ A,
                                                              const double* __restrict__ Z,
                                                              const double* __restrict__ E, double c, int dof,
                                                              int nx, double* __restrict__ Y,
                                                              const int* __restrict__ done) {
  if (done && *done) return;
  __shared__ double red[(NSPLIT - 1) * 2 * 4 * 64];
  const int row0 = blockIdx.x * 16, c0 = blockIdx.y * 16;
  const int lane = threadIdx.x & 63;
  const int cj = c0 + (lane & 15);
  const bool colok = cj < nx;
  const double* ops[2] = {K0, A};
  v4d acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
  if (!split_k_product<2>(ops, dof, Z, nx, colok ? cj : 0, colok, row0, acc, red)) return;
  if (!colok) return;
  const double ex = E[cj];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = row0 + 4 * r + (lane >> 4);
    if (row < dof) {
      const size_t o = (size_t)row * nx + cj;
      Y[o] = Z[o] - c * fma(ex, acc[1][r], acc[0][r]);
    }
  }
}

// Y[:, x] = P[bin(x)] Z[:, x] for the points of every bin: one wavefront per (16-row tile,
// 16-point tile of a bin's points, bin); tiles past a bin's point count exit at once
__global__ __launch_bounds__(64 * NSPLIT) void k_binned_apply(const double* __restrict__ P, int dof,
                                                              const int* __restrict__ grp_start,
                                                              const int* __restrict__ grp_count,
                                                              const int* __restrict__ order,
                                                              const double* __restrict__ Z, int nx,
                                                              double* __restrict__ Y,
                                                              const int* __restrict__ done) {
  if (done && *done) return;
  const int bin = blockIdx.z;
  const int cnt = grp_count[bin];
  const int j0 = blockIdx.y * 16;
  if (j0 >= cnt) return;
  __shared__ double red[(NSPLIT - 1) * 4 * 64];
  const int lane = threadIdx.x & 63;
  const int jj = j0 + (lane & 15);
  const bool colok = jj < cnt;
  const int col = order[grp_start[bin] + (colok ? jj : j0)];
  const double* ops[1] = {P + (size_t)bin * dof * dof};
  v4d acc[1] = {{0.0, 0.0, 0.0, 0.0}};
  const int row0 = blockIdx.x * 16;
  if (!split_k_product<1>(ops, dof, Z, nx, col, colok, row0, acc, red)) return;
  if (!colok) return;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = row0 + 4 * r + (lane >> 4);
    if (row < dof) Y[(size_t)row * nx + col] = acc[0][r];
  }
}

// Counting sort of the points by bin (stable): grp_start[b], grp_count[b], order[]; the
// serial pass runs on LDS copies (nx <= 4096 points, nbins <= 256 bins)
__global__ __launch_bounds__(64) void k_group_bins(const int* __restrict__ bin, int nx, int nbins,
                                                   int* __restrict__ grp_start, int* __restrict__ grp_count,
                                                   int* __restrict__ order) {
  __shared__ int sb[4096], so[4096], cnt[256], st[256];
  for (int x = threadIdx.x; x < nx; x += 64) sb[x] = bin[x];
  for (int b = threadIdx.x; b < nbins; b += 64) cnt[b] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int x = 0; x < nx; ++x) cnt[sb[x]] += 1;
    int s = 0;
    for (int b = 0; b < nbins; ++b) {
      st[b] = s;
      s += cnt[b];
      cnt[b] = 0;
    }
    for (int x = 0; x < nx; ++x) {
      const int b = sb[x];
      so[st[b] + cnt[b]] = x;
      cnt[b] += 1;
    }
  }
  __syncthreads();
  for (int x = threadIdx.x; x < nx; x += 64) order[x] = so[x];
  for (int b = threadIdx.x; b < nbins; b += 64) {
    grp_start[b] = st[b];
    grp_count[b] = cnt[b];
  }
}

// Arnoldi step i of a GMRES cycle (H row-major, H[r][i] at H[r * ldh + i], rows 0..i+1 of
// column i): rotations 0..i-1 applied to the new column, rotation i formed, g[i], g[i+1]
// updated.  When |g[i+1]| <= tol[0] the cycle has converged: done = 1, kdone = i + 1.  A
// zero subdiagonal (happy breakdown) also ends the cycle.
__global__ void k_gmres_givens(double* __restrict__ H, int ldh, int i, double* __restrict__ cs,
                               double* __restrict__ sn, double* __restrict__ g, const double* __restrict__ tol,
                               int* __restrict__ done, int* __restrict__ kdone) {
  if (threadIdx.x != 0 || blockIdx.x != 0 || *done) return;
  double* h = H + i;                      // H[r][i] at h[r * ldh]
  for (int k = 0; k < i; ++k) {
    const double a = h[k * ldh], b = h[(k + 1) * ldh];
    h[k * ldh] = cs[k] * a + sn[k] * b;
    h[(k + 1) * ldh] = -sn[k] * a + cs[k] * b;
  }
  const double a = h[i * ldh], b = h[(i + 1) * ldh];
  const double r = hypot(a, b);
  const double c = (r > 0.0) ? a / r : 1.0, s = (r > 0.0) ? b / r : 0.0;
  cs[i] = c;
  sn[i] = s;
  h[i * ldh] = r;
  h[(i + 1) * ldh] = 0.0;
  const double gi = g[i];
  g[i] = c * gi;
  g[i + 1] = -s * gi;
  if (fabs(g[i + 1]) <= tol[0] || b == 0.0) {
    *done = 1;
    *kdone = i + 1;
  }
}

// Start of a cycle: beta = |r| (in scal[0]) -> g = (beta, 0, ...), tol[0] = max(atol, rtol |b|)
// with |b| in scal[1], done = 0, kdone = m
__global__ void k_gmres_cycle_start(const double* __restrict__ scal, double atol, double rtol, int m,
                                    double* __restrict__ g, double* __restrict__ tol, int* __restrict__ done,
                                    int* __restrict__ kdone) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int k = 0; k <= m; ++k) g[k] = 0.0;
  g[0] = scal[0];
  tol[0] = fmax(atol, rtol * scal[1]);
  *done = (scal[0] < tol[0]) ? 1 : 0;
  *kdone = (scal[0] < tol[0]) ? 0 : m;
}

// y = R^-1 g for the leading k x k upper triangle of H (rotated), k = *kdone
__global__ void k_gmres_back_solve(const double* __restrict__ H, int ldh, const double* __restrict__ g,
                                   const int* __restrict__ kdone, double* __restrict__ y, int m) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int k = *kdone;
  for (int j = k; j < m; ++j) y[j] = 0.0;
  for (int r = k - 1; r >= 0; --r) {
    double s = g[r];
    for (int j = r + 1; j < k; ++j) s -= H[r * ldh + j] * y[j];
    y[r] = s / H[r * ldh + r];
  }
}

inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }


// C = A B, all row-major and contiguous: A (m, k), B (k, n), C (m, n), on the FP64 matrix
// cores, one wavefront per 16 x 16 tile of C (so a product with a few hundred rows and
// thousands of columns fills the chip), v_mfma_f64_16x16x4f64 over k in steps of 4.  Operand
// layout of the instruction (measured on gfx950): A lane = i + 16 kk, B lane = j + 16 kk;
// element q of the accumulator of lane l is D[4 q + l / 16][l % 16].  Rows, columns and k
// beyond the ends are clamped on load (k padded with zero A entries) and not stored.
typedef double bz_gd4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(64) void k_dgemm_nn(const double* __restrict__ A, int m, int k,
                                                 const double* __restrict__ B, int n, double* __restrict__ C) {
  const int lane = threadIdx.x;
  const int i16 = lane & 15, k4 = lane >> 4;
  const int r0 = blockIdx.y * 16, c0 = blockIdx.x * 16;
  const int row = min(r0 + i16, m - 1), col = min(c0 + i16, n - 1);
  const double* Ar = A + (size_t)row * k;
  bz_gd4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int k0 = 0; k0 < k; k0 += 4) {
    const int kk = k0 + k4;
    const int kc = min(kk, k - 1);
    const double a = kk < k ? Ar[kc] : 0.0;
    const double bv = B[(size_t)kc * n + col];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc, 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int rr = r0 + 4 * q + k4, cc = c0 + i16;
    if (rr < m && cc < n) C[(size_t)rr * n + cc] = acc[q];
  }
}

}  // namespace

extern "C" {

int bz_vstep_matvec(const double* K0, const double* A, const double* Z, const double* E, double c, int dof, int nx,
                    double* Y, const int* done, hipStream_t stream) {
  if (dof < 1 || nx < 1 || !K0 || !A || !Z || !E || !Y) return BZ_EINVAL;
  hipLaunchKernelGGL(k_vstep_matvec, dim3(nblk(dof, 16), nblk(nx, 16)), dim3(64 * NSPLIT), 0, stream, K0, A, Z, E, c, dof, nx,
                     Y, done);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_binned_apply(const double* P, int dof, int nbins, const int* grp_start, const int* grp_count,
                    const int* order, const double* Z, int nx, double* Y, const int* done, hipStream_t stream) {
  if (dof < 1 || nx < 1 || nbins < 1 || !P || !grp_start || !grp_count || !order || !Z || !Y) return BZ_EINVAL;
  hipLaunchKernelGGL(k_binned_apply, dim3(nblk(dof, 16), nblk(nx, 16), nbins), dim3(64 * NSPLIT), 0, stream, P, dof, grp_start,
                     grp_count, order, Z, nx, Y, done);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_group_bins(const int* bin, int nx, int nbins, int* grp_start, int* grp_count, int* order, hipStream_t stream) {
  if (nx < 1 || nx > 4096 || nbins < 1 || nbins > 256 || !bin || !grp_start || !grp_count || !order) return BZ_EINVAL;
  hipLaunchKernelGGL(k_group_bins, dim3(1), dim3(64), 0, stream, bin, nx, nbins, grp_start, grp_count, order);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_gmres_givens(double* H, int ldh, int i, double* cs, double* sn, double* g, const double* tol, int* done,
                    int* kdone, hipStream_t stream) {
  if (i < 0 || ldh <= i || !H || !cs || !sn || !g || !tol || !done || !kdone) return BZ_EINVAL;
  hipLaunchKernelGGL(k_gmres_givens, dim3(1), dim3(64), 0, stream, H, ldh, i, cs, sn, g, tol, done, kdone);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_gmres_cycle_start(const double* scal, double atol, double rtol, int m, double* g, double* tol, int* done,
                         int* kdone, hipStream_t stream) {
  if (m < 1 || !scal || !g || !tol || !done || !kdone) return BZ_EINVAL;
  hipLaunchKernelGGL(k_gmres_cycle_start, dim3(1), dim3(64), 0, stream, scal, atol, rtol, m, g, tol, done, kdone);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_gmres_back_solve(const double* H, int ldh, const double* g, const int* kdone, double* y, int m,
                        hipStream_t stream) {
  if (m < 1 || ldh < m || !H || !g || !kdone || !y) return BZ_EINVAL;
  hipLaunchKernelGGL(k_gmres_back_solve, dim3(1), dim3(64), 0, stream, H, ldh, g, kdone, y, m);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}


int bz_dgemm_nn(const double* A, int m, int k, const double* B, int n, double* C, hipStream_t stream) {
  if (m < 1 || k < 1 || n < 1 || !A || !B || !C || C == A || C == B) return BZ_EINVAL;
  if ((m + 15) / 16 > 65535) return BZ_EINVAL;
  hipLaunchKernelGGL(k_dgemm_nn, dim3((n + 15) / 16, (m + 15) / 16), dim3(64), 0, stream, A, m, k, B, n, C);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

}  // extern "C"
