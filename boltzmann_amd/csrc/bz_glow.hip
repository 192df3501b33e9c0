// Plasma coupling of the 1D3V glow-discharge step on the MI355X: the moments of the
// electron distribution that drive the ion fluid, and the ion-fluid half step with its
// Poisson field.
//
// Reference: glow1d_boltzmann (BESolver/python/glowdischarge_boltzmann_1d.py)
//   bte_to_fluid (:1050-1076), bte_eedf_normalization (:1021-1030)   -> k_glow_moments and the
//                                                                        MOMENTS phase
//   solve_poisson (:1005-1019), fluid_to_bte (:1078-1086), the field
//   of the fluid step (:2221-2222)                                   -> the field phases
//   rhs_fluid (:1088-1149), rhs_fluid_jacobian (:1151-1218),
//   step_fluid "BE" (:1318-1372), glow1d_utils.newton_solver
//   (glow1d_utils.py:131-207)                                        -> the STEP phase
//
// The ion fluid lives on Np <= 200 Chebyshev points.  Its backward-Euler Jacobian is a dense
// Np x Np matrix that changes with the field every half step, and the reference inverts it.
// One workgroup of 1024 threads holds the matrix in registers (thread (r, g) owns row r and
// the columns g Sc .. g Sc + Sc - 1, g < CG = 1024 / Np, Sc = ceil(Np / CG)) and inverts it
// in place by Gauss-Jordan elimination with partial pivoting.  Pivot rows are chosen, never moved: the inverse comes out with
// permuted rows and columns and is applied through the permutation.  The Newton loop of the
// reference, its residuals (two products with the Chebyshev derivative matrix, which stays
// in L2) and the Poisson fields run in the same launch, so a half step is one launch with no
// host round trip.  The whole step is FP64.
#include "bz_common.h"
#include "bz.h"

namespace {

constexpr int GF_THREADS = 1024;
constexpr int GF_WAVES = GF_THREADS / 64;
constexpr int GF_NMAX = 200;
constexpr int GF_NPAD = 384;       // >= CG ceil(Np / CG) for every Np <= 200: padded vectors
constexpr int BZ_GLOW_TILE_THREADS = 512;   // block size of the tile layouts
constexpr int GF_TILE_RED = 5000;  // >= max(TR, TC) Npad of the tile layouts below (25 x 200 at 8 x 10)

enum { V_NE = 0, V_NI, V_E, V_X, V_RR, V_R0, V_W, V_Z, V_DX, V_DI, V_COUNT };

struct alignas(16) GfShared {
  double vec[V_COUNT][GF_NPAD];
  double red[GF_THREADS];
  double prow[GF_THREADS + 64];
  double fcol[2][GF_NMAX];
  double wsum[GF_WAVES];
  double candv[GF_WAVES];
  int candi[GF_WAVES];
  int perm[GF_NMAX];
  int qinv[GF_NMAX];
  double red2[GF_TILE_RED];      // tile path: per thread-column partial sums (T x Npad)
  double tprow[2][GF_NMAX + 8];  // tile path: pivot row / column, double-buffered by step
  double tfcol[2][GF_NMAX + 8];
  int flag;
  double sc[4];
  int isc[4];
};

struct GfCtx {
  int n, CG, r, cg, t, nw;
  bool act;
};

// y[r] = sum_c M[r][c] x[c] for M given transposed (MT[c n + r] = M[r][c], so the lanes of a
// wave read consecutive entries); x and y in LDS and distinct.
__device__ __forceinline__ void gf_matvec(const GfCtx& q, GfShared& S, const double* __restrict__ MT,
                                          const double* x, double* y) {
  if (q.act) {
    double acc = 0.0;
    for (int c = q.cg; c < q.n; c += q.CG) acc = fma(MT[(size_t)c * q.n + q.r], x[c], acc);
    S.red[q.cg * q.n + q.r] = acc;
  }
  __syncthreads();
  if (q.t < q.n) {
    double s = 0.0;
    for (int g = 0; g < q.CG; ++g) s += S.red[g * q.n + q.t];
    y[q.t] = s;
  }
  __syncthreads();
}

// Euclidean norm of v[0..n), the same value in every thread.
__device__ __forceinline__ double gf_norm(const GfCtx& q, GfShared& S, const double* v) {
  double a = (q.t < q.n) ? v[q.t] * v[q.t] : 0.0;
  for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
  if ((q.t & 63) == 0) S.wsum[q.t >> 6] = a;
  __syncthreads();
  if (q.t == 0) {
    double s = 0.0;
    for (int w = 0; w < q.nw; ++w) s += S.wsum[w];
    S.sc[0] = sqrt(s);
  }
  __syncthreads();
  const double nr = S.sc[0];
  __syncthreads();
  return nr;
}

// Poisson field G rp with rp = -alpha (ni - ne), rp[0] = bc, rp[n-1] = 0 and G = -Dp LpD^-1.
__device__ __forceinline__ void gf_field(const GfCtx& q, GfShared& S, const BzGlowFluid& g, double bc, double* out) {
  if (q.t < q.n) {
    const int t = q.t;
    S.vec[V_W][t] = (t == 0) ? bc : (t == q.n - 1 ? 0.0 : -g.alpha * (S.vec[V_NI][t] - S.vec[V_NE][t]));
  }
  __syncthreads();
  gf_matvec(q, S, g.GT, S.vec[V_W], out);
}

// rr = x - h (ki n0 ne - Dp flux),  flux = mu ni1 E - Di Dp ni1 (weak flux mu ni1 E at both
// ends), ni1 = ni + x.
__device__ __forceinline__ void gf_residual(const GfCtx& q, GfShared& S, const BzGlowFluid& g, double h,
                                            const double* x, double* rr) {
  const int t = q.t;
  if (t < q.n) S.vec[V_W][t] = S.vec[V_NI][t] + x[t];
  __syncthreads();
  gf_matvec(q, S, g.DpT, S.vec[V_W], S.vec[V_Z]);
  if (t < q.n) {
    const double ni1 = S.vec[V_W][t];
    const double adv = g.mu_i * ni1 * S.vec[V_E][t];
    S.vec[V_Z][t] = (t == 0 || t == q.n - 1) ? adv : adv - g.Di * S.vec[V_Z][t];
  }
  __syncthreads();
  gf_matvec(q, S, g.DpT, S.vec[V_Z], S.vec[V_W]);
  if (t < q.n) rr[t] = x[t] - h * (g.ki[t] * g.n0 * S.vec[V_NE][t] - S.vec[V_W][t]);
  __syncthreads();
}

// dx[i] = (J^-1 rr)[i] from the permuted in-place inverse X: row rho of X is row qinv[rho] of
// J^-1 with its columns taken in pivot order, (J^-1)[i][perm[m]] = X[perm[i]][m].
template <int SMAX>
__device__ __forceinline__ void gf_apply_inverse(const GfCtx& q, GfShared& S, const double (&X)[SMAX], int Sc,
                                                 const double* rr, double* dx) {
  if (q.t < q.n) S.vec[V_W][q.t] = rr[S.perm[q.t]];
  __syncthreads();
  if (q.act) {
    double acc = 0.0;
    const double* w = &S.vec[V_W][q.cg * Sc];
#pragma unroll
    for (int s = 0; s < SMAX; ++s)
      if (s < Sc) acc = fma(X[s], w[s], acc);     // padding columns hold X = 0, w = 0
    S.red[q.cg * q.n + q.r] = acc;
  }
  __syncthreads();
  if (q.t < q.n) {
    double s = 0.0;
    for (int g = 0; g < q.CG; ++g) s += S.red[g * q.n + q.t];
    dx[S.qinv[q.t]] = s;
  }
  __syncthreads();
}

// In-place Gauss-Jordan inverse of the register-resident matrix.  Thread (r, g) keeps the
// columns g Sc + s of row r; every step all threads rotate their registers by one, so the
// column eliminated at step m (column m % SMAX of group m / SMAX) always sits in register 0
// and the loop body indexes the registers with constants only.  After CG SMAX steps the
// registers are back in place.  Steps whose column is padding only rotate.
template <int SMAX>
__device__ __forceinline__ int gf_gauss_jordan(const GfCtx& q, GfShared& S, double (&X)[SMAX], int Sc) {
  const int n = q.n, t = q.t;
  bool done = !q.act;
  int buf = 0, singular = 0;
  for (int m = 0; m < q.CG * SMAX; ++m) {
    const int cgk = m / SMAX, sk = m - cgk * SMAX;
    const int k = cgk * Sc + sk;
    if (sk >= Sc || k >= n) {
      const double x0 = X[0];
#pragma unroll
      for (int s = 0; s + 1 < SMAX; ++s) X[s] = X[s + 1];
      X[SMAX - 1] = x0;
      continue;
    }
    // (a) column k into LDS, pivot candidates per wave (largest |value|, then lowest row)
    const int w = t >> 6;
    const int lo = cgk * n, hi = cgk * n + n - 1;
    if ((64 * w <= hi) && (64 * w + 63 >= lo)) {
      double a = -1.0;
      int ai = 0x7fffffff;
      if (q.act && q.cg == cgk) {
        const double f = X[0];
        S.fcol[buf][q.r] = f;
        if (!done) {
          a = fabs(f);
          if (!(a >= 0.0)) a = 0.0;
          ai = q.r;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(a, off, 64);
        const int oi = __shfl_xor(ai, off, 64);
        if (ov > a || (ov == a && oi < ai)) {
          a = ov;
          ai = oi;
        }
      }
      if ((t & 63) == 0) {
        S.candv[w] = a;
        S.candi[w] = ai;
      }
    }
    __syncthreads();
    // (b) pivot row p, scaled by 1/piv, into LDS (its column-k entry becomes 1/piv)
    const int wlo = lo >> 6, whi = hi >> 6;
    double bv = S.candv[wlo];
    int p = S.candi[wlo];
    for (int ww = wlo + 1; ww <= whi; ++ww) {
      const double ov = S.candv[ww];
      const int oi = S.candi[ww];
      if (ov > bv || (ov == bv && oi < p)) {
        bv = ov;
        p = oi;
      }
    }
    const double piv = S.fcol[buf][p];
    const double inv = 1.0 / piv;
    if (t == 0) {
      S.perm[k] = p;
      S.qinv[p] = k;
    }
    if (!(piv != 0.0)) singular = 1;
    const bool is_p = q.act && q.r == p;
    const bool own_k = q.cg == cgk;
    if (is_p) {
      double* pw = &S.prow[q.cg * SMAX];
      pw[0] = own_k ? inv : X[0] * inv;
#pragma unroll
      for (int s = 1; s < SMAX; ++s) pw[s] = X[s] * inv;
      done = true;
    }
    __syncthreads();
    // (c) X[r][c] = X[r][c] - f_r prow[c] with X[r][k] taken as 0 first, written one register
    // down (the rotation); the pivot row then takes prow itself
    if (q.act) {
      const double f = S.fcol[buf][q.r];
      const double2* pr = reinterpret_cast<const double2*>(&S.prow[q.cg * SMAX]);
      const double x0 = own_k ? 0.0 : X[0];
      double2 cur = pr[0];
      const double p0 = cur.x;
#pragma unroll
      for (int s = 0; s + 1 < SMAX; ++s) {
        if (((s + 1) & 1) == 0) cur = pr[(s + 1) >> 1];
        const double pv = ((s + 1) & 1) ? cur.y : cur.x;
        X[s] = fma(-f, pv, X[s + 1]);
        // keep only a few prow loads in flight: the matrix already fills most registers
        if ((s & 7) == 6) __builtin_amdgcn_sched_barrier(0);
      }
      X[SMAX - 1] = fma(-f, p0, x0);
      if (is_p) {
        const double* pw = &S.prow[q.cg * SMAX];
#pragma unroll
        for (int s = 0; s + 1 < SMAX; ++s) X[s] = pw[s + 1];
        X[SMAX - 1] = pw[0];
      }
    } else {
      const double x0 = X[0];
#pragma unroll
      for (int s = 0; s + 1 < SMAX; ++s) X[s] = X[s + 1];
      X[SMAX - 1] = x0;
    }
    buf ^= 1;
  }
  __syncthreads();
  return singular;
}

// Newton with the frozen inverse and the damping restarts of glow1d_utils.newton_solver
// (glow1d_utils.py:131-207): apply(rr, dx) writes dx = J^-1 rr.  Writes ni and the status.
template <typename Apply>
__device__ __forceinline__ void gf_newton(const GfCtx& q, GfShared& S, const BzGlowFluid& g, double h, int singular,
                                          Apply apply) {
  const int n = q.n, t = q.t;
  double* x = S.vec[V_X];
  double* rr = S.vec[V_RR];
  double* r0 = S.vec[V_R0];
  if (t < n) x[t] = 0.0;
  __syncthreads();
  gf_residual(q, S, g, h, x, r0);
  const double nr0 = gf_norm(q, S, r0);
  double nr = nr0;
  bool conv = (nr / nr0 < g.rtol) || (nr < g.atol);
  double alpha = 1.0;
  int count = 0;
  while (alpha > 1e-6) {
    count = 0;
    if (t < n) {
      x[t] = 0.0;
      rr[t] = r0[t];
    }
    __syncthreads();
    while (!conv && count < g.max_iter) {
      apply(rr, S.vec[V_DX]);
      if (t < n) x[t] = x[t] - alpha * S.vec[V_DX][t];
      __syncthreads();
      gf_residual(q, S, g, h, x, rr);
      nr = gf_norm(q, S, rr);
      ++count;
      conv = (nr / nr0 < g.rtol) || (nr < g.atol);
      if (isnan(nr) || nr > nr0 * 1e2) break;
    }
    if (!conv) alpha *= 0.25;
    else break;
  }
  if (t < n && conv) {
    const double ni1 = S.vec[V_NI][t] + x[t];
    S.vec[V_NI][t] = ni1;
    g.u[3 * t + 1] = ni1;
  }
  if (t == 0) {
    g.status[0] = conv ? 1 : 0;
    g.status[1] = count;
    if (!conv) g.status[2] += 1;
    g.status[3] = (g.status[3] & 1) | singular;   // bit 0 latches a zero pivot, bit 1 the last path
    g.info[0] = nr;
    g.info[1] = nr / nr0;
    g.info[2] = alpha;
  }
  __syncthreads();
}

// Tile layout of the Jacobian for the pivot-free path: thread (tr, tc), tr < TR, tc < TC,
// holds the BR x BC tile of rows tr BR + i and columns tc BC + j of the matrix padded with
// identity rows and columns to Npad, the multiple of lcm(BR, BC) at or above n.
constexpr int gf_gcd(int a, int b) { return b == 0 ? a : gf_gcd(b, a % b); }

struct GfTile {
  int TR, TC, tr, tc, Npad;
  bool act;
};

template <int BR, int BC>
__device__ __forceinline__ GfTile gf_tile_ctx(const GfCtx& q) {
  constexpr int L = BR / gf_gcd(BR, BC) * BC;      // lcm: both rotations are back in place
  GfTile tl;
  tl.Npad = ((q.n + L - 1) / L) * L;
  tl.TR = tl.Npad / BR;
  tl.TC = tl.Npad / BC;
  tl.act = q.t < tl.TR * tl.TC;
  tl.tr = tl.act ? q.t / tl.TC : 0;
  tl.tc = tl.act ? q.t % tl.TC : 0;
  return tl;
}

// J[r][c] = delta_rc + h (mu_i Dp[r][c] E[c] - Di (Dp Dp_mod)[r][c]) into the tiles, and
// whether J is column diagonally dominant (|J_cc| > sum_{r != c} |J_rc| for every c).  For
// such a matrix partial pivoting never exchanges rows, so elimination in the natural order
// takes the same pivots as the reference's LAPACK inverse.
template <int BR, int BC>
__device__ __forceinline__ bool gf_tile_build(const GfCtx& q, GfShared& S, const BzGlowFluid& g, double h,
                                              const GfTile& tl, double (&X)[BR][BC]) {
  const int n = q.n;
#pragma unroll
  for (int i = 0; i < BR; ++i)
#pragma unroll
    for (int j = 0; j < BC; ++j) {
      const int r = tl.tr * BR + i, c = tl.tc * BC + j;
      const bool in = tl.act && r < n && c < n;
      double dp, dpp;
      if (g.DpTile) {                               // tile-ordered copies: coalesced
        const size_t o = (size_t)(i * BC + j) * BZ_GLOW_TILE_THREADS + q.t;
        dp = g.DpTile[o];
        dpp = g.DppTile[o];
      } else {
        const size_t o = (size_t)(in ? c : 0) * n + (in ? r : 0);
        dp = g.DpT[o];
        dpp = g.DppT[o];
      }
      const double v = (r == c ? 1.0 : 0.0) + h * (g.mu_i * dp * S.vec[V_E][in ? c : 0] - g.Di * dpp);
      X[i][j] = in ? v : (r == c ? 1.0 : 0.0);
    }
  // off-diagonal column sums per thread row, then per column
  if (tl.act) {
#pragma unroll
    for (int j = 0; j < BC; ++j) {
      const int c = tl.tc * BC + j;
      double sabs = 0.0;
#pragma unroll
      for (int i = 0; i < BR; ++i) {
        const int r = tl.tr * BR + i;
        if (r != c) sabs += fabs(X[i][j]);
        else {
          S.tprow[0][c] = fabs(X[i][j]);
          S.vec[V_DI][c] = 1.0 / X[i][j];              // the Jacobi path's D^-1
        }
      }
      S.red2[tl.tr * tl.Npad + c] = sabs;
    }
  }
  if (q.t == 0) S.flag = 1;
  __syncthreads();
  if (q.t < tl.Npad) {
    double sum = 0.0;
    for (int r = 0; r < tl.TR; ++r) sum += S.red2[r * tl.Npad + q.t];
    if (!(S.tprow[0][q.t] > sum)) S.flag = 0;
    S.red[q.t] = sum / S.tprow[0][q.t];            // column ratio (padding columns: 0)
  }
  __syncthreads();
  if (q.t == 0) {
    double rmax = 0.0;
    for (int c = 0; c < tl.Npad; ++c) rmax = fmax(rmax, S.red[c]);
    S.sc[0] = rmax;
  }
  __syncthreads();
  const bool dom = S.flag != 0;
  __syncthreads();
  return dom;
}

template <int BR, int BC>
__device__ __forceinline__ void gf_tile_apply(const GfCtx& q, GfShared& S, const GfTile& tl,
                                              const double (&X)[BR][BC], const double* rr, double* dx);

// dx = J^-1 rr by the Jacobi iteration dx <- dx + D^-1 (rr - J dx) from dx = D^-1 rr, on the
// tile-resident J.  With rho the largest column ratio sum_{r != c} |J_rc| / |J_cc|, the
// iteration matrix D^-1 (J - D) has norm rho in |D x|_1, and the start is already within rho
// of the solution, so after K steps the error is below rho^(K+1) of |D dx|_1; K is chosen for
// 2^-53.  Each step is one tile product and an update, two barriers, against one barrier
// per elimination step (and Npad steps) for the Gauss-Jordan inverse.
template <int BR, int BC>
__device__ __forceinline__ void gf_tile_jacobi(const GfCtx& q, GfShared& S, const GfTile& tl,
                                               const double (&X)[BR][BC], int K, const double* rr, double* dx) {
  const int t = q.t;
  const double* di = S.vec[V_DI];
  if (t < q.n) dx[t] = di[t] * rr[t];
  __syncthreads();
  for (int k = 0; k < K; ++k) {
    // tile products of J dx into per-column-tile partial sums, then each row sums them and
    // updates its dx entry: two barriers a step
    if (tl.act) {
#pragma unroll
      for (int i = 0; i < BR; ++i) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < BC; ++j) acc = fma(X[i][j], dx[tl.tc * BC + j], acc);
        S.red2[tl.tc * tl.Npad + tl.tr * BR + i] = acc;
      }
    }
    __syncthreads();
    if (t < q.n) {
      double w = 0.0;
      for (int c = 0; c < tl.TC; ++c) w += S.red2[c * tl.Npad + t];
      dx[t] = fma(di[t], rr[t] - w, dx[t]);
    }
    __syncthreads();
  }
}

// In-place Gauss-Jordan inverse without row exchanges on the tiles.  All threads rotate their
// tiles by one row (period BR) and one column (period BC) every step, so the pivot row of
// step k is tile row 0 of thread row k / BR and the pivot column tile column 0 of thread
// column k / BC: the loop body uses constant register indices, the pivot row and column go
// to LDS as BC / BR values per thread (indexed in the rotated frame all threads share), one
// barrier a step.  After Npad steps every tile is back in place.
template <int BR, int BC>
__device__ __forceinline__ int gf_tile_gauss_jordan(const GfCtx& q, GfShared& S, const GfTile& tl,
                                                    double (&X)[BR][BC]) {
  int singular = 0;
  if (tl.act && tl.tr == 0)
#pragma unroll
    for (int j = 0; j < BC; ++j) S.tprow[0][tl.tc * BC + j] = X[0][j];
  if (tl.act && tl.tc == 0)
#pragma unroll
    for (int i = 0; i < BR; ++i) S.tfcol[0][tl.tr * BR + i] = X[i][0];
  __syncthreads();
  for (int k = 0; k < tl.Npad; ++k) {
    const int dr = k / BR, dc = k / BC, buf = k & 1;
    const double piv = S.tprow[buf][dc * BC];
    const double inv = 1.0 / piv;
    if (!(piv != 0.0)) singular = 1;
    if (tl.act) {
      const bool pr = tl.tr == dr, pc = tl.tc == dc;
      double qv[BC];
#pragma unroll
      for (int j = 0; j < BC; ++j) qv[j] = S.tprow[buf][tl.tc * BC + j] * inv;
      if (pc) qv[0] = inv;
      const double* fc = &S.tfcol[buf][tl.tr * BR];
      double row0[BC];
#pragma unroll
      for (int j = 0; j < BC; ++j) row0[j] = X[0][j];
      // new(i, j) = base(i, j) - f_i q_j (the pivot row takes q: f = -1 on a zero base; the
      // pivot column -f / piv: zero base), stored one row and one column down; in program
      // order X[i][j] is written after X[i+1][j+1] is read, except the wrap to row 0
#pragma unroll
      for (int i = 0; i < BR; ++i) {
        const int si = (i + 1) % BR;
        const double fi = (pr && si == 0) ? -1.0 : fc[si];
#pragma unroll
        for (int j = 0; j < BC; ++j) {
          const int sj = (j + 1) % BC;
          double base = (si == 0) ? row0[sj] : X[si][sj];
          if ((pr && si == 0) || (pc && sj == 0)) base = 0.0;
          X[i][j] = fma(-fi, qv[sj], base);
        }
      }
    }
    const int k1 = k + 1;
    if (k1 < tl.Npad && tl.act) {
      const int nb = buf ^ 1;
      if (tl.tr == k1 / BR)
#pragma unroll
        for (int j = 0; j < BC; ++j) S.tprow[nb][tl.tc * BC + j] = X[0][j];
      if (tl.tc == k1 / BC)
#pragma unroll
        for (int i = 0; i < BR; ++i) S.tfcol[nb][tl.tr * BR + i] = X[i][0];
    }
    __syncthreads();
  }
  return singular;
}

// dx = X rr for the tile-resident inverse (no permutation on this path).
template <int BR, int BC>
__device__ __forceinline__ void gf_tile_apply(const GfCtx& q, GfShared& S, const GfTile& tl,
                                              const double (&X)[BR][BC], const double* rr, double* dx) {
  if (tl.act) {
#pragma unroll
    for (int i = 0; i < BR; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < BC; ++j) acc = fma(X[i][j], rr[tl.tc * BC + j], acc);
      S.red2[tl.tc * tl.Npad + tl.tr * BR + i] = acc;
    }
  }
  __syncthreads();
  if (q.t < q.n) {
    double s = 0.0;
    for (int c = 0; c < tl.TC; ++c) s += S.red2[c * tl.Npad + q.t];
    dx[q.t] = s;
  }
  __syncthreads();
}

template <int NT, int SMAX, int BR, int BC>
__global__ __launch_bounds__(NT) void k_glow_fluid(BzGlowFluid g, int flags, double h, double bc_before,
                                                           double bc_after) {
  __shared__ GfShared S;
  GfCtx q;
  q.n = g.n;
  q.t = threadIdx.x;
  q.CG = NT / q.n;
  q.nw = NT / 64;
  q.r = q.t % q.n;
  q.cg = q.t / q.n;
  q.act = q.cg < q.CG;
  const int n = q.n, t = q.t;

  for (int i = t; i < V_COUNT * GF_NPAD; i += NT) (&S.vec[0][0])[i] = 0.0;
  __syncthreads();
  if (t < n) {
    S.vec[V_NE][t] = g.u[3 * t + 0];
    S.vec[V_NI][t] = g.u[3 * t + 1];
  }
  __syncthreads();

  if (flags & BZ_GLOW_MOMENTS) {
    // bte_to_fluid: ne = m_mass, Te = m_temp / ne, ki = m_rate / m_mass * rate_scale
    double m[3];
    for (int k = 0; k < 3; ++k) {
      if (q.act) {
        double acc = 0.0;
        for (int sp = q.cg; sp < g.nsplit; sp += q.CG) acc += g.partial[((size_t)sp * 3 + k) * n + q.r];
        S.red[q.cg * n + q.r] = acc;
      }
      __syncthreads();
      m[k] = 0.0;
      if (t < n)
        for (int gg = 0; gg < q.CG; ++gg) m[k] += S.red[gg * n + t];
      __syncthreads();
    }
    if (t < n) {
      const double ne = m[0];
      g.u[3 * t + 0] = ne;
      g.u[3 * t + 2] = m[1] / ne;
      g.ki[t] = m[2] / ne * g.rate_scale;
      S.vec[V_NE][t] = ne;
    }
    __syncthreads();
  }

  if (flags & BZ_GLOW_STEP) {
    gf_field(q, S, g, bc_before, S.vec[V_E]);
    if (t < n) g.E[t] = S.vec[V_E][t];

    bool done = false;
    if constexpr (BR > 0) {
      // pivot-free tile path when J is column diagonally dominant
      const GfTile tl = gf_tile_ctx<BR, BC>(q);
      double Xt[BR][BC];
      if (gf_tile_build<BR, BC>(q, S, g, h, tl, Xt)) {
        const double rho = S.sc[0];
        if (rho <= 0.5) {
          // strongly dominant (the glow's half steps: rho about 0.14): Jacobi iterations on J
          const int K = rho > 0.0 ? (int)ceil(36.7368005696771 / -log(rho)) + 1 : 1;   // 53 ln 2
          gf_newton(q, S, g, h, 4,                                   // bit 2: Jacobi path
                    [&](const double* rr, double* dx) { gf_tile_jacobi<BR, BC>(q, S, tl, Xt, K, rr, dx); });
        } else {
          const int singular = gf_tile_gauss_jordan<BR, BC>(q, S, tl, Xt);
          gf_newton(q, S, g, h, singular,
                    [&](const double* rr, double* dx) { gf_tile_apply<BR, BC>(q, S, tl, Xt, rr, dx); });
        }
        done = true;
      }
    }
    if (!done) {
      // J[r][c] = delta_rc + h (mu_i Dp[r][c] E[c] - Di (Dp Dp_mod)[r][c]), row layout,
      // Gauss-Jordan with partial pivoting
      double X[SMAX];
      const int Sc = (n + q.CG - 1) / q.CG;
#pragma unroll
      for (int s = 0; s < SMAX; ++s) {
        X[s] = 0.0;
        if (s < Sc) {
          const int c = q.cg * Sc + s;
          const bool in = q.act && c < n;
          const size_t o = (size_t)(in ? c : 0) * n + q.r;     // padding lanes read a valid entry
          const double v = (q.r == c ? 1.0 : 0.0) + h * (g.mu_i * g.DpT[o] * S.vec[V_E][c] - g.Di * g.DppT[o]);
          X[s] = in ? v : 0.0;
        }
      }
      const int singular = gf_gauss_jordan<SMAX>(q, S, X, Sc) | 2;     // bit 1: pivoted path
      gf_newton(q, S, g, h, singular,
                [&](const double* rr, double* dx) { gf_apply_inverse<SMAX>(q, S, X, Sc, rr, dx); });
    }
  }

  if (flags & BZ_GLOW_FIELD_AFTER) {
    gf_field(q, S, g, bc_after, S.vec[V_DX]);
    if (t < n) g.bsE[t] = S.vec[V_DX][t] * g.field_scale;
  }
}

// Partial ordinate moments: partial[sp][m][x] = sum over the rows of split sp of
// W[m][row] v[row][x]; 64 points x 4 row lanes per workgroup, grid (ceil(nx/64), nsplit).
__global__ __launch_bounds__(256) void k_glow_moments(const double* __restrict__ v, int rows, int nx,
                                                      const double* __restrict__ W, int nsplit,
                                                      double* __restrict__ partial) {
  const int lane = threadIdx.x & 63, sub = threadIdx.x >> 6;
  const int x = blockIdx.x * 64 + lane;
  const int sp = blockIdx.y;
  const int per = (rows + nsplit - 1) / nsplit;
  const int r0 = sp * per, r1 = min(rows, r0 + per);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  if (x < nx) {
    for (int row = r0 + sub; row < r1; row += 4) {
      const double val = v[(size_t)row * nx + x];
      a0 = fma(W[row], val, a0);
      a1 = fma(W[rows + row], val, a1);
      a2 = fma(W[2 * rows + row], val, a2);
    }
  }
  __shared__ double red[3][4][64];
  red[0][sub][lane] = a0;
  red[1][sub][lane] = a1;
  red[2][sub][lane] = a2;
  __syncthreads();
  if (sub == 0 && x < nx) {
    for (int m = 0; m < 3; ++m)
      partial[((size_t)sp * 3 + m) * nx + x] = red[m][0][lane] + red[m][1][lane] + red[m][2][lane] + red[m][3][lane];
  }
}

}  // namespace

// The kernel variant (register layout) bz_glow_fluid_step uses for n points: 0, 1: row
// layouts of 2 and 8 columns per thread at 1024 threads; 2: 6 x 6 tiles at 512; 3: 8 x 10.
static int gf_variant(int n) {
  const int CG = GF_THREADS / n, Sc = (n + CG - 1) / CG;
  if (Sc <= 2) return 0;
  if (Sc <= 8) return 1;
  return n <= 128 ? 2 : 3;
}

template <int BR, int BC>
__global__ __launch_bounds__(256) void k_glow_tile_order(const double* __restrict__ DpT,
                                                         const double* __restrict__ DppT, int n,
                                                         double* __restrict__ DpTile, double* __restrict__ DppTile) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= BR * BC * BZ_GLOW_TILE_THREADS) return;
  const int t = idx % BZ_GLOW_TILE_THREADS, e = idx / BZ_GLOW_TILE_THREADS;
  const int i = e / BC, j = e % BC;
  constexpr int L = BR / gf_gcd(BR, BC) * BC;
  const int Npad = ((n + L - 1) / L) * L, TC = Npad / BC, TR = Npad / BR;
  double a = 0.0, b = 0.0;
  if (t < TR * TC) {
    const int r = (t / TC) * BR + i, c = (t % TC) * BC + j;
    if (r < n && c < n) {
      a = DpT[(size_t)c * n + r];
      b = DppT[(size_t)c * n + r];
    }
  }
  DpTile[idx] = a;
  DppTile[idx] = b;
}

extern "C" int bz_glow_tile_tables(int n, const double* DpT, const double* DppT, double* DpTile, double* DppTile,
                                   hipStream_t stream) {
  if (n < 3 || n > GF_NMAX || !DpT || !DppT || !DpTile || !DppTile) return BZ_EINVAL;
  const int v = gf_variant(n);
  if (v == 2) {
    hipLaunchKernelGGL((k_glow_tile_order<6, 6>), dim3((36 * BZ_GLOW_TILE_THREADS + 255) / 256), dim3(256), 0,
                       stream, DpT, DppT, n, DpTile, DppTile);
  } else if (v == 3) {
    hipLaunchKernelGGL((k_glow_tile_order<8, 10>), dim3((80 * BZ_GLOW_TILE_THREADS + 255) / 256), dim3(256), 0,
                       stream, DpT, DppT, n, DpTile, DppTile);
  } else {
    return BZ_OK;
  }
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

extern "C" int bz_glow_moments(const double* v, int rows, int nx, const double* W, int nsplit, double* partial,
                               hipStream_t stream) {
  if (!v || !W || !partial || rows <= 0 || nx <= 0 || nsplit <= 0 || nsplit > 65535) return BZ_EINVAL;
  dim3 grid((nx + 63) / 64, nsplit);
  hipLaunchKernelGGL(k_glow_moments, grid, dim3(256), 0, stream, v, rows, nx, W, nsplit, partial);
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}

extern "C" int bz_glow_fluid_step(const BzGlowFluid* g, int flags, double h, double bc_before, double bc_after,
                                  hipStream_t stream) {
  if (!g || g->n < 3 || g->n > GF_NMAX || !g->u) return BZ_EINVAL;
  if ((flags & ~(BZ_GLOW_MOMENTS | BZ_GLOW_STEP | BZ_GLOW_FIELD_AFTER)) != 0) return BZ_EINVAL;
  if ((flags & BZ_GLOW_MOMENTS) && (!g->partial || g->nsplit <= 0 || !g->ki)) return BZ_EINVAL;
  if ((flags & BZ_GLOW_STEP) && (!g->DpT || !g->DppT || !g->GT || !g->ki || !g->E || !g->status || !g->info ||
                                 g->max_iter < 0))
    return BZ_EINVAL;
  if ((flags & BZ_GLOW_FIELD_AFTER) && (!g->GT || !g->bsE)) return BZ_EINVAL;
  // tile layouts: 8 x 10 for 128 < n <= 200 (Npad <= 200, 25 x 20 threads), 6 x 6 for
  // 90 < n <= 128 (Npad <= 132, 22 x 22 threads): LDS reductions of max(TR, TC) Npad entries
  static_assert(25 * 200 <= GF_TILE_RED && 22 * 132 <= GF_TILE_RED, "tile reduction buffer");
  static_assert(25 * 20 <= 512 && 22 * 22 <= 512, "tile threads");
  const BzGlowFluid gv = *g;
  // register layout: the matrix fits in registers either at 1024 threads (4 waves per SIMD,
  // 128 registers each) with up to 8 columns per thread, or at 512 threads (256 registers)
  // with up to 32 (n <= 128) or 100 columns per thread
  switch (gf_variant(g->n)) {
    case 0:
      hipLaunchKernelGGL((k_glow_fluid<1024, 2, 0, 0>), dim3(1), dim3(1024), 0, stream, gv, flags, h, bc_before, bc_after);
      break;
    case 1:
      hipLaunchKernelGGL((k_glow_fluid<1024, 8, 0, 0>), dim3(1), dim3(1024), 0, stream, gv, flags, h, bc_before, bc_after);
      break;
    case 2:
      hipLaunchKernelGGL((k_glow_fluid<512, 32, 6, 6>), dim3(1), dim3(512), 0, stream, gv, flags, h, bc_before, bc_after);
      break;
    default:
      hipLaunchKernelGGL((k_glow_fluid<512, 100, 8, 10>), dim3(1), dim3(512), 0, stream, gv, flags, h, bc_before,
                         bc_after);
  }
  BZ_LAUNCH_CHECK();
  return BZ_OK;
}
