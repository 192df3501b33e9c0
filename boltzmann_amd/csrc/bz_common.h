// Shared device helpers for the velocity-space kernels (gfx950, FP64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BZ_OK 0
#define BZ_EINVAL -1
#define BZ_EHIP -2
#define BZ_ENOSPACE -3

#define BZ_HIP_CHECK(expr)                           \
  do {                                               \
    hipError_t _e = (expr);                          \
    if (_e != hipSuccess) return BZ_EHIP;            \
  } while (0)

#define BZ_LAUNCH_CHECK()                            \
  do {                                               \
    hipError_t _e = hipGetLastError();               \
    if (_e != hipSuccess) return BZ_EHIP;            \
  } while (0)

// Physical constants are never hard-coded in device code: the host passes the
// derived scalars (vth, c_gamma, kappa, ...) computed from scipy.constants.
namespace bz {

// Knot interval l with t[l] <= x < t[l+1], l in [p, np_r-1]; x equal to the last
// knot belongs to the last interval.  Returns -1 outside [t[0], t[np_r+p]].
__device__ __forceinline__ int find_interval(const double* __restrict__ t, int npr, int p, double x) {
  const double lo = t[0], hi = t[npr + p];
  if (!(x >= lo) || !(x <= hi)) return -1;
  if (x == hi) return npr - 1;
  int a = p, b = npr - 1;        // t[a] <= x < t[b+1]
  while (a < b) {
    int m = (a + b + 1) >> 1;
    if (t[m] <= x) a = m; else b = m - 1;
  }
  return a;
}

// Cubic B-spline values N[r] = B_{l-3+r}(x) and derivatives D[r] = B'_{l-3+r}(x)
// on interval l (de Boor triangular recursion; derivative from the order-2 values).
__device__ __forceinline__ void bspline3_eval(const double* __restrict__ t, int l, double x,
                                              double N[4], double D[4]) {
  double left[4], right[4];
  double n2[3] = {0.0, 0.0, 0.0};
  N[0] = 1.0; N[1] = N[2] = N[3] = 0.0;
  for (int j = 1; j <= 3; ++j) {
    left[j] = x - t[l + 1 - j];
    right[j] = t[l + j] - x;
    double saved = 0.0;
    for (int r = 0; r < j; ++r) {
      double temp = N[r] / (right[r + 1] + left[j - r]);
      N[r] = saved + right[r + 1] * temp;
      saved = left[j - r] * temp;
    }
    N[j] = saved;
    if (j == 2) { n2[0] = N[0]; n2[1] = N[1]; n2[2] = N[2]; }
  }
  // B'_{i,3} = 3 B_{i,2}/(t[i+3]-t[i]) - 3 B_{i+1,2}/(t[i+4]-t[i+1]),  i = l-3+r
  // order-2 values available: B_{l-2,2}, B_{l-1,2}, B_{l,2} = n2[0..2]
  for (int r = 0; r < 4; ++r) {
    const int i = l - 3 + r;
    const double bi = (r >= 1) ? n2[r - 1] : 0.0;     // B_{i,2}
    const double bi1 = (r <= 2) ? n2[r] : 0.0;        // B_{i+1,2}
    const double d1 = t[i + 3] - t[i];
    const double d2 = t[i + 4] - t[i + 1];
    double v = 0.0;
    if (d1 != 0.0) v += 3.0 * bi / d1;
    if (d2 != 0.0) v -= 3.0 * bi1 / d2;
    D[r] = v;
  }
}

// Value (dorder 0) or derivative (dorder 1) of B_k at x; zero outside the support.
__device__ __forceinline__ double bspline3_single(const double* __restrict__ t, int npr, int k, double x, int dorder) {
  int l = find_interval(t, npr, 3, x);
  if (l < 0) return 0.0;
  const int r = k - (l - 3);
  if (r < 0 || r > 3) return 0.0;
  double N[4], D[4];
  bspline3_eval(t, l, x, N, D);
  return dorder ? D[r] : N[r];
}

// B_k(x - dv) - B_k(x) (0 < dv << x) as minus the integral of B_k' over [x - dv, x], split at
// the knots inside and with 2 Gauss-Legendre points per piece (exact: B_k' is quadratic on
// each knot interval).  The piece widths are formed from dv, not from the rounded end point
// x - dv, so the result keeps full relative precision; the difference of the two values
// keeps only ~eps x / dv of it (the elastic post-collision speed v sqrt(1 - 2m/M):
// dv / v ~ 1.4e-5).
__device__ __forceinline__ double bspline3_diff_down(const double* __restrict__ t, int npr, int k, double x,
                                                     double dv) {
  int l = find_interval(t, npr, 3, x - dv);
  if (l < 0 || !(dv > 0.0)) return 0.0;
  const double g = 0.28867513459481287;          // 1 / (2 sqrt(3))
  double s = 0.0, off = dv;                      // left end of the current piece: x - off
  while (off > 0.0 && l <= npr - 1) {
    const double roff = fmax(0.0, x - t[l + 1]); // right end of the piece: x - roff
    if (roff < off) {
      const double h = off - roff, m = x - 0.5 * (off + roff);
      const int r = k - (l - 3);
      if (r >= 0 && r <= 3) {
        double N[4], D1[4], D2[4];
        bspline3_eval(t, l, m - g * h, N, D1);
        bspline3_eval(t, l, m + g * h, N, D2);
        s += 0.5 * h * (D1[r] + D2[r]);
      }
      off = roff;
    }
    ++l;
  }
  return -s;
}

// Piecewise-linear table with flat extrapolation (interval chosen like
// numpy searchsorted(side='left') clipped to [1, len-1]).
__device__ __forceinline__ double interp_linear_flat(const double* __restrict__ e, const double* __restrict__ s,
                                                     int len, double x) {
  if (x < e[0]) return s[0];
  if (x > e[len - 1]) return s[len - 1];
  int a = 0, b = len;            // first index with e[idx] >= x
  while (a < b) {
    int m = (a + b) >> 1;
    if (e[m] < x) a = m + 1; else b = m;
  }
  int idx = a < 1 ? 1 : (a > len - 1 ? len - 1 : a);
  const int lo = idx - 1, hi = idx;
  const double slope = (s[hi] - s[lo]) / (e[hi] - e[lo]);
  return slope * (x - e[lo]) + s[lo];
}

// s + c += a * b with the rounding errors of the product (fma) and of the sum (Knuth's
// TwoSum) collected in c: the result is as accurate as a sum in twice the precision.
__device__ __forceinline__ void two_sum_acc(double a, double b, double& s, double& c) {
#pragma clang fp contract(off)   // a contracted s + a*b would lose the product's rounding error
  const double p = a * b;
  const double pe = fma(a, b, -p);
  const double t = s + p;
  const double z = t - s;
  const double se = (s - (t - z)) + (p - z);
  s = t;
  c += se + pe;
}

// 64-lane sum (result broadcast to every lane of the wave).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace bz
