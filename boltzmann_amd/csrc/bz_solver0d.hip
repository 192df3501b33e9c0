// Fused batched 0D steady-state Newton solver, FP64, one 64-lane wavefront per instance.
//
// Reference: bte_0d3v_batched.steady_state_solve / get_rhs_and_jacobian
// (BESolver/python/bte_0d3v_batched.py:602-821, 882-979).  The reference forms the
// dense Jacobian J = Q^T L Q - mu I in the (n-1)-dim basis Q of u-perp and inverts it.
// Here the same Newton step is obtained from the bordered system in the original
// B-spline basis (L = M^-1 K):
//      (K - mu M) y = c M uhat + d M x,   uhat^T y = uhat^T x,   d = w.y - mu
// (d = 0 reproduces the reference's Jacobian, which omits the derivative of the
// growth term; d free gives the exact Jacobian).  K - mu M is sparse: a narrow
// band (|offset| <= bw) plus the inelastic gain band.  Ordering the radial index
// from high to low energy puts the gain band strictly below the diagonal, so LU
// without pivoting keeps U inside the narrow band and the fill stays in L, which is
// never stored: the forward substitution runs during the factorisation.
//
// Data layout (all vectors in the permuted order i' = (np_r-1-p)*ns + l):
//   * shared per grid: operator values in row profile storage (columns
//     first_col[i]..min(i+bw, n-1)), nop arrays: C_0..C_{S-1}, C_Tg, A_E, M
//   * per instance: K profile (sum of the scaled operators, built once per solve),
//     U rows (W doubles each, 1/U_jj first), scalar results.
// Lane/slot mapping: each lane owns R row slots; a host-computed schedule assigns
// rows to slots so that no two rows alive at the same step share a slot (a row is
// alive from step first_col-2W, when its prefetch starts, to its pivot step).
#include "bz_common.h"
#include "../../include/bz.h"

namespace {

using namespace bz;

constexpr int kMaxOps = 16;

struct GridDev {
  int n, ns, npr, bw, S, nop, P, nslots;
  const int* first_col;     // [n]
  const int* prof_off;      // [n+1]
  const double* opprof;     // [nop][P]
  const int* slot_first;    // [nslots]
  const int* sched_next;    // [n]
  const double* muhat;      // [n]  M uhat
  const double* uhat;       // [n]
  const double* u;          // [n]
  const double* wsp;        // [nop-1][n]  W_m = Op_m^T M^-1 u (zero rows: not in the growth term)
  const double* mchol;      // [n][4]  banded Cholesky of M (permuted), L[i, i-d*ns]
};

struct BatchDev {
  int B;
  const double* n0;         // [B]
  const double* Tg;         // [B]
  const double* E;          // [B]
  const double* ns_by_n0;   // [S][B]
  double* x;                // (n, B) original order, in/out
  double* kprof;            // [B][P] scratch
  double* urows;            // [B][n][WP] scratch
  double* vec;              // [B][4][n] scratch (A, Bv, P, WV)
  int* iters;               // [B]
  int* status;              // [B]  0 converged, 1 max_iter, 2 line search failed
  double* res_abs;          // [B]
  double* res0;             // [B]
  double atol, rtol, alpha_min, alpha_rho, ls_c;
  int max_iter, mode;       // mode 0: reference Jacobian (d = 0), 1: exact Jacobian
  double tg_scale;          // coefficient of C_Tg is n0 * Tg * tg_scale
};

__device__ __forceinline__ int orig_index(int i, int ns, int npr) {
  const int q = i / ns, l = i - q * ns;
  return (npr - 1 - q) * ns + l;
}

__device__ __forceinline__ double fast_rcp(double a) {
  double r = __builtin_amdgcn_rcp(a);
  r = fma(r, fma(-a, r, 1.0), r);
  r = fma(r, fma(-a, r, 1.0), r);
  return r;
}

// y = K x  (profile storage of this instance), result into y (LDS or global), lane-strided rows
__device__ void profile_matvec(const GridDev& g, const double* __restrict__ kp, const int* __restrict__ fcol,
                               const int* __restrict__ poff, const double* __restrict__ x, double* __restrict__ y,
                               int lane) {
  for (int i = lane; i < g.n; i += 64) {
    const int f = fcol[i];
    const int o = poff[i];
    const int len = poff[i + 1] - o;
    double s = 0.0;
    for (int e = 0; e < len; ++e) s = fma(kp[o + e], x[f + e], s);
    y[i] = s;
  }
}

// In-place solve M z = b in the permuted order (per l-block banded Cholesky,
// half bandwidth 3 in block index, stride ns); lanes 0..ns-1 each own one block.
__device__ void mass_solve(const GridDev& g, double* __restrict__ b, int lane) {
  const int ns = g.ns, npr = g.npr;
  if (lane < ns) {
    const int l = lane;
    for (int q = 0; q < npr; ++q) {
      const int i = q * ns + l;
      double s = b[i];
      for (int d = 1; d <= 3 && q - d >= 0; ++d) s -= g.mchol[i * 4 + d] * b[i - d * ns];
      b[i] = s / g.mchol[i * 4];
    }
    for (int q = npr - 1; q >= 0; --q) {
      const int i = q * ns + l;
      double s = b[i];
      for (int d = 1; d <= 3 && q + d < npr; ++d) {
        const int k = i + d * ns;
        s -= g.mchol[k * 4 + d] * b[k];
      }
      b[i] = s / g.mchol[i * 4];
    }
  }
}

__device__ __forceinline__ double dot_lds(const double* __restrict__ a, const double* __restrict__ b, int n, int lane) {
  double s = 0.0;
  for (int i = lane; i < n; i += 64) s = fma(a[i], b[i], s);
  return wave_sum(s);
}

// Norm of the projected residual r = (I - uhat uhat^T) (A + al*Bv - mu_a (X + al*P)).
__device__ double proj_res_norm(const double* A, const double* Bv, const double* X, const double* P,
                                const double* __restrict__ uhat, int n, double al, double mua, int lane) {
  double su = 0.0, ss = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double R = (A[i] + al * Bv[i]) - mua * (X[i] + al * P[i]);
    su = fma(uhat[i], R, su);
    ss = fma(R, R, ss);
  }
  su = wave_sum(su);
  // second pass for the explicit projection (avoids cancellation in |R|^2 - (u.R)^2)
  double s2 = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double R = (A[i] + al * Bv[i]) - mua * (X[i] + al * P[i]);
    const double rp = R - su * uhat[i];
    s2 = fma(rp, rp, s2);
  }
  (void)ss;
  return sqrt(wave_sum(s2));
}

// Factorise A = K - mu M (profile, permuted order, no pivoting) and forward-eliminate
// the two right-hand sides b1 = M uhat, b2 = M x.  Writes U rows (1/U_jj first) to
// urows and the eliminated right-hand sides to Y1, Y2 (LDS, indexed by pivot row).
template <int W, int R>
__device__ void factor_forward(const GridDev& g, const double* __restrict__ kprof, double mu,
                               const int* __restrict__ s_fcol, const int* __restrict__ s_poff,
                               const int* __restrict__ s_next, const double* __restrict__ s_muhat,
                               const double* __restrict__ s_mx, double* __restrict__ Y1, double* __restrict__ Y2,
                               double* __restrict__ s_piv, double* __restrict__ urows, int WP, int lane) {
  const int n = g.n, bw = g.bw;
  const double* __restrict__ mprof = g.opprof + (size_t)(g.nop - 1) * g.P;
  int row[R], fc[R], lc[R], off[R];
  double w[R][W], nk[R][W], nm[R][W], b1[R], b2[R];

#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int rw = g.slot_first[lane * R + r];
    row[r] = rw;
    if (rw >= 0) {
      fc[r] = s_fcol[rw];
      lc[r] = min(rw + bw, n - 1);
      off[r] = s_poff[rw] - fc[r];
      b1[r] = s_muhat[rw];
      b2[r] = s_mx[rw];
    } else {
      fc[r] = 0x7fffffff; lc[r] = -1; off[r] = 0; b1[r] = 0.0; b2[r] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const int c0 = k, c1 = W + k;
      double v0 = 0.0, a1 = 0.0, m1 = 0.0;
      if (c0 >= fc[r] && c0 <= lc[r]) v0 = kprof[off[r] + c0] - mu * mprof[off[r] + c0];
      if (c1 >= fc[r] && c1 <= lc[r]) { a1 = kprof[off[r] + c1]; m1 = mprof[off[r] + c1]; }
      w[r][k] = v0; nk[r][k] = a1; nm[r][k] = m1;
    }
  }

  for (int j0 = 0; j0 < n; j0 += W) {
#pragma unroll
    for (int t = 0; t < W; ++t) {
      const int j = j0 + t;
      if (j >= n) break;
      // --- locate the pivot row j
      int hit = -1;
#pragma unroll
      for (int r = 0; r < R; ++r) hit = (row[r] == j) ? r : hit;
      const unsigned long long bal = __ballot(hit >= 0);
      const int lp = __builtin_amdgcn_readfirstlane(__ffsll((long long)bal) - 1);
      // --- pivot lane publishes its window (column order j..j+W-1) and rhs
      if (lane == lp) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (hit == r) {
#pragma unroll
            for (int k = 0; k < W; ++k) s_piv[k] = w[r][(t + k) % W];
            s_piv[W] = b1[r];
            s_piv[W + 1] = b2[r];
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      double pv[W];
#pragma unroll
      for (int k = 0; k < W; ++k) pv[k] = s_piv[k];
      const double pb1 = s_piv[W], pb2 = s_piv[W + 1];
      const double rinv = fast_rcp(pv[0]);
      // --- store U row j and the eliminated rhs
      {
        double v = rinv;
#pragma unroll
        for (int k = 1; k < W; ++k) v = (lane == k) ? pv[k] : v;
        if (lane < W) urows[(size_t)j * WP + lane] = v;
        if (lane == 0) { Y1[j] = pb1; Y2[j] = pb2; }
      }
      // --- eliminate column j, shift in column j+W, prefetch column j+2W
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double m = w[r][t] * rinv;
#pragma unroll
        for (int k = 1; k < W; ++k) w[r][(t + k) % W] = fma(-m, pv[k], w[r][(t + k) % W]);
        b1[r] = fma(-m, pb1, b1[r]);
        b2[r] = fma(-m, pb2, b2[r]);
        w[r][t] = fma(-mu, nm[r][t], nk[r][t]);
      }
      // --- retire the pivot row; its slot takes the next scheduled row
      if (lane == lp) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (hit == r) {
            const int nr = s_next[row[r]];
            row[r] = nr;
#pragma unroll
            for (int k = 0; k < W; ++k) { w[r][k] = 0.0; nk[r][k] = 0.0; nm[r][k] = 0.0; }
            if (nr >= 0) {
              fc[r] = s_fcol[nr];
              lc[r] = min(nr + bw, n - 1);
              off[r] = s_poff[nr] - fc[r];
              b1[r] = s_muhat[nr];
              b2[r] = s_mx[nr];
            } else {
              fc[r] = 0x7fffffff; lc[r] = -1; off[r] = 0; b1[r] = 0.0; b2[r] = 0.0;
            }
          }
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int c = j + 2 * W;
        double a = 0.0, m2 = 0.0;
        if (c >= fc[r] && c <= lc[r]) { a = kprof[off[r] + c]; m2 = mprof[off[r] + c]; }
        nk[r][t] = a;
        nm[r][t] = m2;
      }
    }
  }
}

// Back substitution U z = y for the two right-hand sides held in Y1/Y2 (in place).
// Lane 0 solves Y1, lane 1 solves Y2; U rows are read as broadcasts.
template <int W>
__device__ void back_substitute(int n, const double* __restrict__ urows, int WP, double* __restrict__ Y1,
                                double* __restrict__ Y2, int lane) {
  if (lane < 2) {
    double* Y = lane == 0 ? Y1 : Y2;
    double z[W];
#pragma unroll
    for (int k = 0; k < W; ++k) z[k] = 0.0;
    const int jtop = ((n + W - 1) / W) * W;
    for (int j0 = jtop - W; j0 >= 0; j0 -= W) {
#pragma unroll
      for (int tt = W - 1; tt >= 0; --tt) {
        const int j = j0 + tt;
        if (j < n) {
          const double* ur = urows + (size_t)j * WP;
          double s = Y[j];
#pragma unroll
          for (int k = W - 1; k >= 1; --k) s = fma(-ur[k], z[(tt + k) % W], s);
          const double zj = s * ur[0];
          z[tt] = zj;
          Y[j] = zj;
        }
      }
    }
  }
}

template <int W, int R>
__global__ __launch_bounds__(64) void k_newton0d(GridDev g, BatchDev b) {
  extern __shared__ double lds[];
  const int n = g.n, lane = threadIdx.x, inst = blockIdx.x;
  const int WP = (W + 1) & ~1;
  double* X = lds;
  double* Z1 = X + n;
  double* Z2 = Z1 + n;
  double* MXv = Z2 + n;
  double* MUH = MXv + n;
  double* PIV = MUH + n;                                   // W + 2 (padded to 32)
  int* FC = reinterpret_cast<int*>(PIV + 32);
  int* PO = FC + n;
  int* NX = PO + (n + 1);

  double* A = b.vec + (size_t)inst * 4 * n;                // M^-1 K x
  double* Bv = A + n;                                      // M^-1 K p
  double* Pv = Bv + n;                                     // Newton step
  double* WV = Pv + n;                                     // growth vector w (mu = w.x)
  double* kprof = b.kprof + (size_t)inst * g.P;
  double* urows = b.urows + (size_t)inst * n * WP;

  // ---------------- per-grid tables into LDS
  for (int i = lane; i < n; i += 64) {
    FC[i] = g.first_col[i];
    NX[i] = g.sched_next[i];
    MUH[i] = g.muhat[i];
  }
  for (int i = lane; i <= n; i += 64) PO[i] = g.prof_off[i];

  // ---------------- per-instance coefficients and K profile
  const int S = g.S;
  double coef[kMaxOps];
  const double n0 = b.n0[inst];
#pragma unroll
  for (int m = 0; m < kMaxOps; ++m) coef[m] = 0.0;
  for (int s = 0; s < S; ++s) coef[s] = n0 * b.ns_by_n0[(size_t)s * b.B + inst];
  coef[S] = n0 * b.Tg[inst] * b.tg_scale;
  coef[S + 1] = b.E[inst];
  const int nk = g.nop - 1;
  for (int e = lane; e < g.P; e += 64) {
    double v = 0.0;
    for (int m = 0; m < nk; ++m) v = fma(coef[m], g.opprof[(size_t)m * g.P + e], v);
    kprof[e] = v;
  }
  for (int i = lane; i < n; i += 64) {
    double v = 0.0;
    for (int m = 0; m < nk; ++m) v = fma(coef[m], g.wsp[(size_t)m * n + i], v);
    WV[i] = v;
  }
  // initial iterate (original order, (n, B) layout) -> permuted, normalised u.x = 1
  for (int i = lane; i < n; i += 64) X[i] = b.x[(size_t)orig_index(i, g.ns, g.npr) * b.B + inst];
  __syncthreads();
  {
    const double ux = dot_lds(g.u, X, n, lane);
    for (int i = lane; i < n; i += 64) X[i] = X[i] / ux;
  }
  __syncthreads();

  // ---------------- initial residual
  profile_matvec(g, kprof, FC, PO, X, A, lane);
  __syncthreads();
  mass_solve(g, A, lane);
  __syncthreads();
  double mu = dot_lds(WV, X, n, lane);
  for (int i = lane; i < n; i += 64) { Pv[i] = 0.0; Bv[i] = 0.0; }
  __syncthreads();
  const double r0 = proj_res_norm(A, Bv, X, Pv, g.uhat, n, 0.0, mu, lane);
  double rcur = r0;
  double abs_err = r0;
  const double tol = b.atol + b.rtol * r0;
  int it = 0, status = 1;

  while (it < b.max_iter) {
    if (b.mode == 0) { if (!(abs_err > tol)) { status = 0; break; } abs_err = rcur; }
    else { if (!(rcur > tol)) { status = 0; break; } }
    mu = dot_lds(WV, X, n, lane);
    // M x -> MXv : rows of M are the last operator's profile
    {
      const double* mp = g.opprof + (size_t)(g.nop - 1) * g.P;
      profile_matvec(g, mp, FC, PO, X, MXv, lane);
    }
    __syncthreads();
    factor_forward<W, R>(g, kprof, mu, FC, PO, NX, MUH, MXv, Z1, Z2, PIV, urows, WP, lane);
    __syncthreads();
    back_substitute<W>(n, urows, WP, Z1, Z2, lane);
    __syncthreads();
    // --- 2x2 system for (c, d)
    const double s11 = dot_lds(g.uhat, Z1, n, lane);
    const double s12 = dot_lds(g.uhat, Z2, n, lane);
    const double ux = dot_lds(g.uhat, X, n, lane);
    double c, d;
    if (b.mode == 0) {
      c = ux / s11;
      d = 0.0;
    } else {
      const double s21 = dot_lds(WV, Z1, n, lane);
      const double s22 = dot_lds(WV, Z2, n, lane) - 1.0;
      const double det = s11 * s22 - s12 * s21;
      c = (ux * s22 - s12 * mu) / det;
      d = (s11 * mu - s21 * ux) / det;
    }
    for (int i = lane; i < n; i += 64) Pv[i] = fma(c, Z1[i], fma(d, Z2[i], -X[i]));
    __syncthreads();
    // --- b = M^-1 K p
    profile_matvec(g, kprof, FC, PO, Pv, Bv, lane);
    __syncthreads();
    mass_solve(g, Bv, lane);
    __syncthreads();
    // --- Armijo backtracking on 0.5 |r|^2 (bte_0d3v_batched.py:936-952)
    const double mup = dot_lds(WV, Pv, n, lane);
    const double ae2 = rcur * rcur;
    double alpha = 1.0;
    bool diverged = false;
    double rnew = proj_res_norm(A, Bv, X, Pv, g.uhat, n, alpha, mu + alpha * mup, lane);
    while (0.5 * rnew * rnew > 0.5 * ae2 - b.ls_c * alpha * ae2) {
      alpha *= b.alpha_rho;
      if (alpha < b.alpha_min) { diverged = true; break; }
      rnew = proj_res_norm(A, Bv, X, Pv, g.uhat, n, alpha, mu + alpha * mup, lane);
    }
    if (diverged) { status = 2; break; }
    for (int i = lane; i < n; i += 64) {
      X[i] = fma(alpha, Pv[i], X[i]);
      A[i] = fma(alpha, Bv[i], A[i]);
    }
    __syncthreads();
    rcur = rnew;
    if (b.mode == 0 && (it % 10) == 0) abs_err = rcur;
    ++it;
  }
  if (status == 1 && ((b.mode == 0 && !(abs_err > tol)) || (b.mode != 0 && !(rcur > tol)))) status = 0;

  // ---------------- write back (original order)
  for (int i = lane; i < n; i += 64) b.x[(size_t)orig_index(i, g.ns, g.npr) * b.B + inst] = X[i];
  if (lane == 0) {
    b.iters[inst] = it;
    b.status[inst] = status;
    b.res_abs[inst] = rcur;
    b.res0[inst] = r0;
  }
}

template <int W, int R>
int launch_newton(const GridDev& g, const BatchDev& b, hipStream_t stream) {
  const size_t lds = (size_t)(5 * g.n + 32) * sizeof(double) + (size_t)(3 * g.n + 1) * sizeof(int);
  if (lds > 64 * 1024) return BZ_EINVAL;
  hipLaunchKernelGGL((k_newton0d<W, R>), dim3(b.B), dim3(64), lds, stream, g, b);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

// Gather dense (n x n, original order) operators into permuted row-profile storage.
__global__ void k_gather_profile(const double* __restrict__ dense, int n, int ns, int npr,
                                 const int* __restrict__ first_col, const int* __restrict__ prof_off,
                                 double scale, double* __restrict__ out) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const int o = prof_off[i], len = prof_off[i + 1] - o, f = first_col[i];
  const int oi = orig_index(i, ns, npr);
  for (int e = threadIdx.x; e < len; e += blockDim.x) {
    const int oc = orig_index(f + e, ns, npr);
    out[o + e] = scale * dense[(size_t)oi * n + oc];
  }
}

// QoIs per instance from normalised h (n, B) (bte_0d3v_batched.py:1163-1217).
__global__ void k_qoi0d(int n, int ns, int B, int G, const double* __restrict__ h, const double* __restrict__ mass,
                        const double* __restrict__ temp, const double* __restrict__ mob, const double* __restrict__ dif,
                        const double* __restrict__ rates, const double* __restrict__ E, const double* __restrict__ n0,
                        double mm_scale, double mw0v3, double mw0v5, double c_gamma, int effective,
                        double* __restrict__ hout, double* __restrict__ out) {
  const int inst = blockIdx.x, lane = threadIdx.x;
  if (inst >= B) return;
  const int npr = n / ns;
  double sm = 0.0;
  for (int i = lane; i < n; i += 64) sm = fma(mass[i], h[(size_t)i * B + inst], sm);
  sm = wave_sum(sm);
  const double scale = sm * mm_scale;       // normalised_distribution scale
  double mm = 0.0, tt = 0.0, mo = 0.0, df = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double v = h[(size_t)i * B + inst] / scale;
    hout[(size_t)i * B + inst] = v;
    mm = fma(mass[i], v, mm);
    tt = fma(temp[i], v, tt);
  }
  for (int p = lane; p < npr; p += 64) {
    const double h0 = h[(size_t)(p * ns) * B + inst] / scale;
    df = fma(dif[p], h0, df);
    if (ns > 1) mo = fma(mob[p], sqrt(3.0) * (h[(size_t)(p * ns + 1) * B + inst] / scale), mo);
  }
  mm = wave_sum(mm) * mw0v3;
  tt = wave_sum(tt);
  mo = wave_sum(mo);
  df = wave_sum(df);
  const double energy = 1.5 * tt * mw0v5 / mm;
  const double ef = effective ? (E[inst] / n0[inst]) : (1.0 / n0[inst]);
  double* o = out + (size_t)inst * (3 + G);
  double rr[16];
  for (int gi = 0; gi < G && gi < 16; ++gi) {
    double s = 0.0;
    for (int p = lane; p < npr; p += 64) s = fma(rates[(size_t)gi * npr + p], h[(size_t)(p * ns) * B + inst] / scale, s);
    rr[gi] = wave_sum(s);
  }
  if (lane == 0) {
    o[0] = energy;
    o[1] = mo * (-(c_gamma / (3.0 * ef)));
    o[2] = df * (c_gamma / 3.0);
    for (int gi = 0; gi < G && gi < 16; ++gi) o[3 + gi] = rr[gi];
  }
}

// Small dense helpers for setup: y = alpha * op(A) x  (A row-major m x k).
__global__ void k_gemv(const double* __restrict__ A, int m, int k, int trans, const double* __restrict__ x,
                       double alpha, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int rows = trans ? k : m;
  if (i >= rows) return;
  double s = 0.0;
  if (!trans) for (int c = 0; c < k; ++c) s = fma(A[(size_t)i * k + c], x[c], s);
  else for (int r = 0; r < m; ++r) s = fma(A[(size_t)r * k + i], x[r], s);
  y[i] = alpha * s;
}

// Permute a vector from original to permuted order (or back when inverse != 0).
__global__ void k_permute(const double* __restrict__ in, int n, int ns, int npr, int inverse, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = orig_index(i, ns, npr);
  if (!inverse) out[i] = in[o]; else out[o] = in[i];
}

// Permuted banded Cholesky rows of M from the radial Cholesky of the reversed block.
// lrev[q*4+d] = L_rev[q, q-d]; every l-block shares it.
__global__ void k_mchol_perm(const double* __restrict__ lrev, int ns, int npr, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npr * ns) return;
  const int q = i / ns;
  for (int d = 0; d < 4; ++d) out[i * 4 + d] = lrev[q * 4 + d];
}

}  // namespace

extern "C" {

int bz_gather_profile(const double* dense, int n, int ns, int npr, const int* first_col, const int* prof_off,
                      double scale, double* out, hipStream_t stream) {
  if (n != ns * npr || n < 1) return BZ_EINVAL;
  hipLaunchKernelGGL(k_gather_profile, dim3(n), dim3(64), 0, stream, dense, n, ns, npr, first_col, prof_off, scale,
                     out);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_gemv(const double* A, int m, int k, int trans, const double* x, double alpha, double* y, hipStream_t stream) {
  if (m < 1 || k < 1) return BZ_EINVAL;
  const int rows = trans ? k : m;
  hipLaunchKernelGGL(k_gemv, dim3((rows + 127) / 128), dim3(128), 0, stream, A, m, k, trans, x, alpha, y);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_permute(const double* in, int n, int ns, int npr, int inverse, double* out, hipStream_t stream) {
  if (n != ns * npr || n < 1) return BZ_EINVAL;
  hipLaunchKernelGGL(k_permute, dim3((n + 127) / 128), dim3(128), 0, stream, in, n, ns, npr, inverse, out);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_mchol_perm(const double* lrev, int ns, int npr, double* out, hipStream_t stream) {
  const int n = ns * npr;
  if (n < 1) return BZ_EINVAL;
  hipLaunchKernelGGL(k_mchol_perm, dim3((n + 127) / 128), dim3(128), 0, stream, lrev, ns, npr, out);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

int bz_steady_solve_0d(const bz_grid0d_t* gp, const bz_batch0d_t* bp, hipStream_t stream) {
  if (!gp || !bp) return BZ_EINVAL;
  GridDev g;
  g.n = gp->n; g.ns = gp->ns; g.npr = gp->npr; g.bw = gp->bw; g.S = gp->S; g.nop = gp->nop; g.P = gp->P;
  g.nslots = gp->nslots;
  g.first_col = gp->first_col; g.prof_off = gp->prof_off; g.opprof = gp->opprof; g.slot_first = gp->slot_first;
  g.sched_next = gp->sched_next; g.muhat = gp->muhat; g.uhat = gp->uhat; g.u = gp->u; g.wsp = gp->wsp;
  g.mchol = gp->mchol;
  if (g.n != g.ns * g.npr || g.n < 8 || g.nop != g.S + 3 || g.nop > kMaxOps || g.S < 1) return BZ_EINVAL;
  BatchDev b;
  b.B = bp->B; b.n0 = bp->n0; b.Tg = bp->Tg; b.E = bp->E; b.ns_by_n0 = bp->ns_by_n0; b.x = bp->x;
  b.kprof = bp->kprof; b.urows = bp->urows; b.vec = bp->vec; b.iters = bp->iters; b.status = bp->status;
  b.res_abs = bp->res_abs; b.res0 = bp->res0; b.atol = bp->atol; b.rtol = bp->rtol; b.alpha_min = bp->alpha_min;
  b.alpha_rho = bp->alpha_rho; b.ls_c = bp->ls_c; b.max_iter = bp->max_iter; b.mode = bp->mode;
  b.tg_scale = bp->tg_scale;
  if (b.B < 1 || b.max_iter < 0) return BZ_EINVAL;
  // the backtracking loop terminates only for 0 < alpha_rho < 1 and alpha_min > 0
  if (!(b.alpha_rho > 0.0 && b.alpha_rho < 1.0) || !(b.alpha_min > 0.0)) return BZ_EINVAL;
  if (!(b.atol >= 0.0) || !(b.rtol >= 0.0)) return BZ_EINVAL;
  const int Wneed = g.bw + 1;
  const int R = (g.nslots + 63) / 64;
  if (g.nslots != 64 * R) return BZ_EINVAL;
  if (Wneed <= 8) {
    if (R == 1) return launch_newton<8, 1>(g, b, stream);
    if (R == 2) return launch_newton<8, 2>(g, b, stream);
    if (R == 3) return launch_newton<8, 3>(g, b, stream);
  } else if (Wneed <= 11) {
    if (R == 1) return launch_newton<11, 1>(g, b, stream);
    if (R == 2) return launch_newton<11, 2>(g, b, stream);
    if (R == 3) return launch_newton<11, 3>(g, b, stream);
  } else if (Wneed <= 17) {
    if (R == 2) return launch_newton<17, 2>(g, b, stream);
    if (R == 3) return launch_newton<17, 3>(g, b, stream);
    if (R == 4) return launch_newton<17, 4>(g, b, stream);
  }
  return BZ_EINVAL;
}

int bz_solver0d_window(int bw) {
  const int Wneed = bw + 1;
  if (Wneed <= 8) return 8;
  if (Wneed <= 11) return 11;
  if (Wneed <= 17) return 17;
  return -1;
}

int bz_qoi_0d(int n, int ns, int B, int G, const double* h, const double* mass, const double* temp,
              const double* mob, const double* dif, const double* rates, const double* E, const double* n0,
              double mm_scale, double mw0v3, double mw0v5, double c_gamma, int effective, double* hout, double* out,
              hipStream_t stream) {
  if (n < 1 || ns < 1 || n % ns || B < 1 || G < 0 || G > 16) return BZ_EINVAL;
  hipLaunchKernelGGL(k_qoi0d, dim3(B), dim3(64), 0, stream, n, ns, B, G, h, mass, temp, mob, dif, rates, E, n0,
                     mm_scale, mw0v3, mw0v5, c_gamma, effective, hout, out);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

}  // extern "C"
