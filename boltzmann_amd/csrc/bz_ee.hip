// Electron-electron Fokker-Planck operator in the B-spline x harmonic basis (gfx950, FP64).
//
// Assembly (once per grid):
//   * partial moments of every B-spline at the quadrature points (Rosenbluth potentials,
//     collision_operator_spherical.py:786-909), the potential kernels m1, m2 per l and the
//     potential operators H_l, G_l = M^-1 [int v^2 B_p m(., k)] (:911-970);
//   * the trilinear weak-form tensors cc_a, cc_b (:973-1050) stored banded:
//     cc[p][k-p+3][r-p+3][q][l][s] for |k-p|, |r-p| <= 3 (three cubic B-splines overlap only
//     there), 49 ns^3 values per radial index instead of the dense n^3 tensor.
// Per Newton step of the batched solver (the factored form of bte_0d3v_batched.py:444-476,
// 702-709): potentials h = H x, g = G x of every instance and the Coulomb prefactor
// ne gamma_a(x) (:652-661); the e-e operator at fixed potentials is then added to the
// instance's K profile by k_ee_profile_add (bz_solver0d.hip).  Per Krylov step: the
// potentials of the direction and the potential-derivative part of the Jacobian (k_ee_t2).
#include <algorithm>

#include "bz_common.h"

namespace {

using namespace bz;

// Values, first and second derivatives of the 4 cubic B-splines nonzero on interval l
// (t[l] <= x < t[l+1]): N[r], D1[r], D2[r] for B_{l-3+r}.
__device__ void bspline3_d2(const double* __restrict__ t, int l, double x, double N[4], double D1[4], double D2[4]) {
  // order 1 (linear) on interval l: B_{l-1,1}, B_{l,1}
  const double h0 = t[l + 1] - t[l];
  const double b1[2] = {(t[l + 1] - x) / h0, (x - t[l]) / h0};
  // derivatives of order-2 B_{j,2}, j = l-2, l-1, l: B'_{j,2} = 2 (B_{j,1}/(t[j+2]-t[j]) - B_{j+1,1}/(t[j+3]-t[j+1]))
  double d2o[3];
  for (int a = 0; a < 3; ++a) {
    const int j = l - 2 + a;
    const double bj = (a >= 1) ? b1[a - 1] : 0.0;      // B_{j,1}
    const double bj1 = (a <= 1) ? b1[a] : 0.0;         // B_{j+1,1}
    const double e1 = t[j + 2] - t[j], e2 = t[j + 3] - t[j + 1];
    double v = 0.0;
    if (e1 != 0.0) v += 2.0 * bj / e1;
    if (e2 != 0.0) v -= 2.0 * bj1 / e2;
    d2o[a] = v;
  }
  bspline3_eval(t, l, x, N, D1);
  for (int r = 0; r < 4; ++r) {
    const int i = l - 3 + r;
    const double di = (r >= 1) ? d2o[r - 1] : 0.0;      // B'_{i,2}
    const double di1 = (r <= 2) ? d2o[r] : 0.0;         // B'_{i+1,2}
    const double e1 = t[i + 3] - t[i], e2 = t[i + 4] - t[i + 1];
    double v = 0.0;
    if (e1 != 0.0) v += 3.0 * di / e1;
    if (e2 != 0.0) v -= 3.0 * di1 / e2;
    D2[r] = v;
  }
}

// x^m rounded once: the product carried as hi + lo (TwoProduct by fma), the reciprocal of
// a negative power corrected by one Newton step on hi + lo
__device__ __forceinline__ double ipow_acc(double x, int m) {
#pragma clang fp contract(off)
  double hi = 1.0, lo = 0.0;
  const int e = m < 0 ? -m : m;
  for (int k = 0; k < e; ++k) {
    const double p = hi * x;
    const double pe = fma(hi, x, -p) + lo * x;
    hi = p + pe;
    lo = pe - (hi - p);
  }
  if (m >= 0) return hi + lo;
  const double q = 1.0 / hi;
  return q + q * (fma(-hi, q, 1.0) - lo * q);
}

// sum over the knot pieces of supp(B_i) clipped to [xb, xe] of int v^m B_i, qn Gauss points
// per piece (integrate_xm_splines, collision_operator_spherical.py:803-817), summed in
// compensated arithmetic
__device__ double piece_integral(const double* __restrict__ t, int npr, int i, int m, double xb, double xe,
                                 const double* __restrict__ qx, const double* __restrict__ qw, int qn) {
  double s = 0.0, e = 0.0;          // compensated sum over all pieces and points
  for (int kn = i + 1; kn <= i + 4; ++kn) {
    const double a = fmax(xb, t[kn - 1]), b = fmin(xe, t[kn]);
    if (b > a) {
      const double hw = 0.5 * (b - a), c = 0.5 * (b + a);
      const int l = kn - 1;                       // knot interval [t[l], t[l+1]]
      for (int g = 0; g < qn; ++g) {
        const double x = hw * qx[g] + c;
        double N[4], D[4];
        bspline3_eval(t, l, x, N, D);
        const int r = i - (l - 3);
        two_sum_acc(hw * qw[g], ipow_acc(x, m) * N[r], s, e);
      }
    }
  }
  return s + e;
}

// kind 0: P[i, j] = int_{t_i}^{min(x_j, t_(i+4))} v^m B_i  (0 below the support)
// kind 1: Q[i, j] = int_{max(x_j, t_i)}^{t_(i+4)} v^-m B_i (0 above the support)
__global__ void k_partial_moment(const double* __restrict__ t, int npr, const double* __restrict__ gx, int nq, int m,
                                 int kind, const double* __restrict__ qx, const double* __restrict__ qw, int qn,
                                 double* __restrict__ out) {
  const long total = (long)npr * nq;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e / nq), j = (int)(e % nq);
    const double lo = t[i], hi = t[i + 4], x = gx[j];
    double v = 0.0;
    if (kind == 0) {
      if (x >= lo) v = piece_integral(t, npr, i, m, lo, fmin(x, hi), qx, qw, qn);
    } else {
      if (x <= hi) v = piece_integral(t, npr, i, -m, fmax(x, lo), hi, qx, qw, qn);
    }
    out[e] = v;
  }
}

// m1, m2 of harmonic l at the points (collision_operator_spherical.py:915-920)
__global__ void k_potential_kernels(int l, double m_ab, const double* __restrict__ gx, int nq, int npr,
                                    const double* __restrict__ P2, const double* __restrict__ P4,
                                    const double* __restrict__ Q1, const double* __restrict__ Q3, double* __restrict__ m1,
                                    double* __restrict__ m2) {
  const long total = (long)npr * nq;
  const double f1 = (l - 0.5) / (l + 1.5);
  const double c1 = (2.0 / (2 * l + 1)) * (1.0 + m_ab);
  const double c2 = -2.0 / (4.0 * l * l - 1.0);
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const double x = gx[e % nq];
    m1[e] = c1 * (P2[e] / ipow_acc(x, l + 1) + Q1[e] * ipow_acc(x, l));
    m2[e] = c2 * ((P2[e] / ipow_acc(x, l - 1) - f1 * P4[e] / ipow_acc(x, l + 1)) +
                  (Q3[e] * ipow_acc(x, l) - f1 * Q1[e] * ipow_acc(x, l + 2)));
  }
}

// C[i, j] = alpha * sum_k A[i*sai + k*sak] B[k*sbk + j*sbj]  (small set-up products, summed in
// compensated arithmetic)
__global__ void k_gemm_strided(int m, int n, int kk, double alpha, const double* __restrict__ A, long sai, long sak,
                               const double* __restrict__ B, long sbk, long sbj, double* __restrict__ C, long ldc) {
  const long total = (long)m * n;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    double s = 0.0, c = 0.0;
    for (int k = 0; k < kk; ++k) two_sum_acc(A[i * sai + k * sak], B[k * sbk + j * sbj], s, c);
    C[(long)i * ldc + j] = alpha * (s + c);
  }
}

// Banded trilinear tensors.  One thread per (p, dk, dr): the points of every knot interval
// where B_p, B_k and B_r overlap (q Gauss points per interval), all (q, l, s) combinations.
// Term tables: combination c = (q*ns + l)*ns + s has monomials toff[c] .. toff[c+1]-1 of
// (coef, power of v, derivative orders of B_p, B_k, B_r); kind a and b are separate tables.
__global__ void k_cc_banded(const double* __restrict__ t, int npr, int ns, const double* __restrict__ glx,
                            const double* __restrict__ glw, int q, const int* __restrict__ ta_off,
                            const double* __restrict__ ta_c, const int* __restrict__ ta_i,
                            const int* __restrict__ tb_off, const double* __restrict__ tb_c,
                            const int* __restrict__ tb_i, double* __restrict__ ccA, double* __restrict__ ccB) {
  const int ncomb = ns * ns * ns;
  const long total = (long)npr * 49;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int p = (int)(e / 49), dk = (int)((e / 7) % 7) - 3, dr = (int)(e % 7) - 3;
    const int k = p + dk, r = p + dr;
    double* oa = ccA + e * ncomb;
    double* ob = ccB + e * ncomb;
    if (k < 0 || k >= npr || r < 0 || r >= npr) {
      for (int c = 0; c < ncomb; ++c) { oa[c] = 0.0; ob[c] = 0.0; }
      continue;
    }
    // intervals l with all three supports: l in [max(p,k,r), min(p,k,r)+3]
    const int l0 = max(p, max(k, r)), l1 = min(p, min(k, r)) + 3;
    for (int c = 0; c < ncomb; ++c) {
      // every monomial of every point summed in compensated arithmetic (the terms of one
      // combination have both signs)
      double sa = 0.0, ea = 0.0, sb = 0.0, eb = 0.0;
      for (int l = l0; l <= l1; ++l) {
        if (!(t[l + 1] > t[l])) continue;
        const double hw = 0.5 * (t[l + 1] - t[l]), cx = 0.5 * (t[l + 1] + t[l]);
        for (int g = 0; g < q; ++g) {
          const double x = hw * glx[g] + cx, w = hw * glw[g];
          double N[4], D1[4], D2[4];
          bspline3_d2(t, l, x, N, D1, D2);
          const int rp = p - (l - 3), rk = k - (l - 3), rr = r - (l - 3);
          const double bp[3] = {N[rp], D1[rp], D2[rp]};
          const double bk[3] = {N[rk], D1[rk], D2[rk]};
          const double br[3] = {N[rr], D1[rr], D2[rr]};
          for (int m = ta_off[c]; m < ta_off[c + 1]; ++m)
            two_sum_acc(w * ta_c[m], ipow_acc(x, ta_i[4 * m]) * bp[ta_i[4 * m + 1]] * bk[ta_i[4 * m + 2]] *
                                         br[ta_i[4 * m + 3]], sa, ea);
          for (int m = tb_off[c]; m < tb_off[c + 1]; ++m)
            two_sum_acc(w * tb_c[m], ipow_acc(x, tb_i[4 * m]) * bp[tb_i[4 * m + 1]] * bk[tb_i[4 * m + 2]] *
                                         br[tb_i[4 * m + 3]], sb, eb);
        }
      }
      oa[c] = sa + ea;
      ob[c] = sb + eb;
    }
  }
}

// h[b, r*ns + s] = sum_k H[s][r][k] x[(k*ns + s), b]  (and g with G); x (n, B), h, g (B, n)
// cols = 0: h, g (B, n);  cols = 1: h, g (n, B) like x
__global__ void k_ee_potentials(const double* __restrict__ H, const double* __restrict__ G, int ns, int npr,
                                const double* __restrict__ x, int B, int cols, double* __restrict__ h,
                                double* __restrict__ g) {
  const int n = ns * npr;
  const long total = (long)n * B;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int b = (int)(e % B), row = (int)(e / B);
    const int r = row / ns, s = row % ns;
    const double* Hr = H + ((long)s * npr + r) * npr;
    const double* Gr = G + ((long)s * npr + r) * npr;
    double sh = 0.0, sg = 0.0;
    for (int k = 0; k < npr; ++k) {
      const double xv = x[(long)(k * ns + s) * B + b];
      sh = fma(Hr[k], xv, sh);
      sg = fma(Gr[k], xv, sg);
    }
    const long o = cols ? e : (long)b * n + row;
    h[o] = sh;
    g[o] = sg;
  }
}

// s + c += a * b, product and sum errors collected in c (fma TwoProd, Knuth TwoSum)
__device__ __forceinline__ void dd_acc(double a, double b, double& s, double& c) {
#pragma clang fp contract(off)   // a contracted s + a*b would lose the product's rounding error
  const double p = a * b;
  const double pe = fma(a, b, -p);
  const double t = s + p;
  const double z = t - s;
  const double se = (s - (t - z)) + (p - z);
  s = t;
  c += se + pe;
}

// h, g (B, n) with the potentials to twice FP64 precision: h + h_lo, g + g_lo
__global__ void k_ee_potentials_dd(const double* __restrict__ H, const double* __restrict__ G, int ns, int npr,
                                   const double* __restrict__ x, int B, double* __restrict__ h,
                                   double* __restrict__ g, double* __restrict__ hlo, double* __restrict__ glo) {
#pragma clang fp contract(off)
  const int n = ns * npr;
  const long total = (long)n * B;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int b = (int)(e % B), row = (int)(e / B);
    const int r = row / ns, s = row % ns;
    const double* Hr = H + ((long)s * npr + r) * npr;
    const double* Gr = G + ((long)s * npr + r) * npr;
    double sh = 0.0, ch = 0.0, sg = 0.0, cg = 0.0;
    for (int k = 0; k < npr; ++k) {
      const double xv = x[(long)(k * ns + s) * B + b];
      dd_acc(Hr[k], xv, sh, ch);
      dd_acc(Gr[k], xv, sg, cg);
    }
    const long o = (long)b * n + row;
    const double th = sh + ch, tg = sg + cg;          // renormalise (FastTwoSum)
    h[o] = th;
    hlo[o] = ch - (th - sh);
    g[o] = tg;
    glo[o] = cg - (tg - sg);
  }
}

// The potential-derivative part of the e-e Jacobian (bte_0d3v_batched.py:734-743, the
// cc_12 x term), in M-space:
//   t2[b, (p,q)] = negam[b] sum_{(k,l),(r,s)} (cc_a[(p,q),(k,l),(r,s)] hv[(r,s), b]
//                                              + cc_b[...] gv[(r,s), b]) x[(k,l), b]
// x, hv, gv (n, B); t2 (B, n).  cc banded: |k - p|, |r - p| <= 3.
template <int NS>
__global__ __launch_bounds__(256) void k_ee_t2(int npr, int B, const double* __restrict__ ccA,
                                               const double* __restrict__ ccB, const double* __restrict__ x,
                                               const double* __restrict__ hv, const double* __restrict__ gv,
                                               const double* __restrict__ negam, const int* __restrict__ active,
                                               double* __restrict__ t2) {
  // one workgroup: radial row p, 256 instances; the tensor blocks of row p for one q at a
  // time in LDS (read by every thread), the instance's potentials at r = p-3..p+3 in registers
  constexpr int NC2 = NS * NS, NC = NS * NS * NS;
  __shared__ double SA[49 * NC2], SB[49 * NC2];
  const int p = blockIdx.y, tid = threadIdx.x;
  const int inst = blockIdx.x * 256 + tid;
  const int n = NS * npr;
  const bool ok = inst < B && (!active || active[inst]);
  double h[7][NS], g[7][NS];
#pragma unroll
  for (int d = 0; d < 7; ++d) {
    const int r = p + d - 3;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const bool in = ok && r >= 0 && r < npr;
      h[d][s] = in ? hv[(long)(r * NS + s) * B + inst] : 0.0;
      g[d][s] = in ? gv[(long)(r * NS + s) * B + inst] : 0.0;
    }
  }
  const double ng = ok ? negam[inst] : 0.0;
  for (int q = 0; q < NS; ++q) {
    __syncthreads();
    for (int e = tid; e < 49 * NC2; e += 256) {
      const int dkdr = e / NC2, ls = e - dkdr * NC2;
      const long src = ((long)p * 49 + dkdr) * NC + q * NC2 + ls;
      SA[e] = ccA[src];
      SB[e] = ccB[src];
    }
    __syncthreads();
    if (!ok) continue;
    double sum = 0.0;
#pragma unroll
    for (int dk = 0; dk < 7; ++dk) {
      const int k = p + dk - 3;
      if (k < 0 || k >= npr) continue;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        const double xv = x[(long)(k * NS + l) * B + inst];
        double a = 0.0;
#pragma unroll
        for (int dr = 0; dr < 7; ++dr) {
#pragma unroll
          for (int ss = 0; ss < NS; ++ss) {
            const int e = (dk * 7 + dr) * NC2 + l * NS + ss;
            a = fma(SA[e], h[dr][ss], a);
            a = fma(SB[e], g[dr][ss], a);
          }
        }
        sum = fma(a, xv, sum);
      }
    }
    t2[(long)inst * n + p * NS + q] = ng * sum;
  }
}

// negam[b] = ne[b] * gamma_a(x_b): m0 = c_m0 (mass . x), kT = |c_kt (top . x) / m0|,
// Lambda = c_lam kT^1.5 / sqrt(ne), gamma = ln(Lambda) c_gam  (bte_0d3v_batched.py:652-661)
__global__ void k_ee_gamma(const double* __restrict__ x, int n, int B, const double* __restrict__ mass,
                           const double* __restrict__ top, const double* __restrict__ ne, double c_m0, double c_kt,
                           double c_lam, double c_gam, double* __restrict__ negam) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double sm = 0.0, st = 0.0;
  for (int i = 0; i < n; ++i) {
    const double xv = x[(long)i * B + b];
    sm = fma(mass[i], xv, sm);
    st = fma(top[i], xv, st);
  }
  const double m0 = c_m0 * sm;
  const double kT = fabs(c_kt * st / m0);
  const double lam = c_lam * kT * sqrt(kT) / sqrt(ne[b]);
  negam[b] = ne[b] * log(lam) * c_gam;
}

int grid_for(long n) { return (int)std::min<long>((n + 255) / 256, 16384); }

}  // namespace

extern "C" {

int bz_partial_moment(const double* t, int npr, const double* gx, int nq, int m, int kind, const double* qx,
                      const double* qw, int qn, double* out, hipStream_t stream) {
  if (npr < 4 || nq < 1 || qn < 1 || (kind != 0 && kind != 1) || !t || !gx || !qx || !qw || !out) return BZ_EINVAL;
  hipLaunchKernelGGL(k_partial_moment, dim3(grid_for((long)npr * nq)), dim3(256), 0, stream, t, npr, gx, nq, m, kind, qx,
                     qw, qn, out);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_potential_kernels(int l, double m_ab, const double* gx, int nq, int npr, const double* P2, const double* P4,
                         const double* Q1, const double* Q3, double* m1, double* m2, hipStream_t stream) {
  if (l < 0 || nq < 1 || npr < 1 || !gx || !P2 || !P4 || !Q1 || !Q3 || !m1 || !m2) return BZ_EINVAL;
  hipLaunchKernelGGL(k_potential_kernels, dim3(grid_for((long)npr * nq)), dim3(256), 0, stream, l, m_ab, gx, nq, npr, P2,
                     P4, Q1, Q3, m1, m2);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_gemm_strided(int m, int n, int k, double alpha, const double* A, long sai, long sak, const double* B, long sbk,
                    long sbj, double* C, long ldc, hipStream_t stream) {
  if (m < 1 || n < 1 || k < 1 || !A || !B || !C || ldc < n) return BZ_EINVAL;
  hipLaunchKernelGGL(k_gemm_strided, dim3(grid_for((long)m * n)), dim3(256), 0, stream, m, n, k, alpha, A, sai, sak, B,
                     sbk, sbj, C, ldc);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_cc_banded(const double* t, int npr, int ns, const double* glx, const double* glw, int q, const int* ta_off,
                 const double* ta_c, const int* ta_i, const int* tb_off, const double* tb_c, const int* tb_i,
                 double* ccA, double* ccB, hipStream_t stream) {
  if (npr < 4 || ns < 1 || ns > 8 || q < 1 || !t || !glx || !glw || !ta_off || !ta_c || !ta_i || !tb_off || !tb_c ||
      !tb_i || !ccA || !ccB)
    return BZ_EINVAL;
  hipLaunchKernelGGL(k_cc_banded, dim3(grid_for((long)npr * 49)), dim3(256), 0, stream, t, npr, ns, glx, glw, q, ta_off,
                     ta_c, ta_i, tb_off, tb_c, tb_i, ccA, ccB);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_ee_potentials(const double* H, const double* G, int ns, int npr, const double* x, int B, double* h, double* g,
                     hipStream_t stream) {
  if (ns < 1 || npr < 1 || B < 1 || !H || !G || !x || !h || !g) return BZ_EINVAL;
  hipLaunchKernelGGL(k_ee_potentials, dim3(grid_for((long)ns * npr * B)), dim3(256), 0, stream, H, G, ns, npr, x, B, 0,
                     h, g);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_ee_potentials_dd(const double* H, const double* G, int ns, int npr, const double* x, int B, double* h,
                        double* g, double* h_lo, double* g_lo, hipStream_t stream) {
  if (ns < 1 || npr < 1 || B < 1 || !H || !G || !x || !h || !g || !h_lo || !g_lo) return BZ_EINVAL;
  hipLaunchKernelGGL(k_ee_potentials_dd, dim3(grid_for((long)ns * npr * B)), dim3(256), 0, stream, H, G, ns, npr, x, B,
                     h, g, h_lo, g_lo);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_ee_potentials_cols(const double* H, const double* G, int ns, int npr, const double* x, int B, double* h,
                          double* g, hipStream_t stream) {
  if (ns < 1 || npr < 1 || B < 1 || !H || !G || !x || !h || !g) return BZ_EINVAL;
  hipLaunchKernelGGL(k_ee_potentials, dim3(grid_for((long)ns * npr * B)), dim3(256), 0, stream, H, G, ns, npr, x, B, 1,
                     h, g);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_ee_jacobian_t2(int ns, int npr, int B, const double* ccA, const double* ccB, const double* x, const double* hv,
                      const double* gv, const double* negam, const int* active, double* t2, hipStream_t stream) {
  if (ns < 1 || ns > 5 || npr < 1 || B < 1 || !ccA || !ccB || !x || !hv || !gv || !negam || !t2) return BZ_EINVAL;
  const dim3 grid((B + 255) / 256, npr);
  switch (ns) {
    case 1: hipLaunchKernelGGL(k_ee_t2<1>, grid, dim3(256), 0, stream, npr, B, ccA, ccB, x, hv, gv, negam, active, t2); break;
    case 2: hipLaunchKernelGGL(k_ee_t2<2>, grid, dim3(256), 0, stream, npr, B, ccA, ccB, x, hv, gv, negam, active, t2); break;
    case 3: hipLaunchKernelGGL(k_ee_t2<3>, grid, dim3(256), 0, stream, npr, B, ccA, ccB, x, hv, gv, negam, active, t2); break;
    case 4: hipLaunchKernelGGL(k_ee_t2<4>, grid, dim3(256), 0, stream, npr, B, ccA, ccB, x, hv, gv, negam, active, t2); break;
    default: hipLaunchKernelGGL(k_ee_t2<5>, grid, dim3(256), 0, stream, npr, B, ccA, ccB, x, hv, gv, negam, active, t2);
  }
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_ee_gamma(const double* x, int n, int B, const double* mass, const double* top, const double* ne, double c_m0,
                double c_kt, double c_lam, double c_gam, double* negam, hipStream_t stream) {
  if (n < 1 || B < 1 || !x || !mass || !top || !ne || !negam) return BZ_EINVAL;
  hipLaunchKernelGGL(k_ee_gamma, dim3((B + 255) / 256), dim3(256), 0, stream, x, n, B, mass, top, ne, c_m0, c_kt, c_lam,
                     c_gam, negam);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

}  // extern "C"
